// k_lrq.hip -- Rank-M plus packed-u8 LRQ verifier: encode, materialize_k.
//
// CompactLrqVerifier::build / materialize_k (src/rknn_query.cpp:679-753,
// policy "floor", bits 8) accumulate in binary64 with GCC-contracted fma.
// The kernels carry those doubles as bit patterns through softdouble.hpp, so
// coefficients, codes and thresholds equal the reference's bit for bit while
// the GPU issues no FP64 instruction.
//
// Device layout: codes are kept k-major (nk x n) next to the row-major file
// layout, so materialize_k streams 8 coefficients (32 B) + 1 code byte per row
// and writes one float: 37 B/row, HBM-bound.
#include "common.hpp"
#include "kernels.hpp"

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

// t_k(i) = float(max(0, slack * (mu_k + sum_m c_m G_mk + cb_k[code])))
__global__ void k_lrq_materialize(const float* __restrict__ coeff, const uint8_t* __restrict__ codes_km,
                                  uint64_t n, uint32_t nk, uint32_t rankm,
                                  const float* __restrict__ mu, const float* __restrict__ basis,
                                  const float* __restrict__ cb, uint32_t code_size, uint32_t k,
                                  uint64_t slack_bits, float* __restrict__ thr) {
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
    thr[i] = sd_lt_zero_or_zero(scaled) ? 0.0f : sd_to_float(scaled);
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
                                  uint32_t k_one_based, uint64_t slack_bits, float* thr,
                                  hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lrq_materialize, dim3(blocks_for(n, 256)), dim3(256), 0, s, coeff,
                       codes_kmajor, n, nk, rankm, mu, basis, cb, code_size, k_one_based - 1,
                       slack_bits, thr);
    return hipGetLastError();
}

hipError_t launch_transpose_codes(const uint8_t* codes, uint64_t n, uint32_t nk,
                                  uint8_t* codes_kmajor, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_transpose_codes, dim3(blocks_for(n * nk, 256)), dim3(256), 0, s, codes, n,
                       nk, codes_kmajor);
    return hipGetLastError();
}
