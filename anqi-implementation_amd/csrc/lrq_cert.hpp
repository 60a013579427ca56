// lrq_cert.hpp -- the verifier's per-row encode and per-element threshold
// decode in FP32 float-float with certified rounding (dfloat.hpp), each
// element that cannot be certified recomputed in soft binary64
// (softdouble.hpp).  Values are bit-identical to CompactLrqVerifier::build /
// materialize_k (src/rknn_query.cpp:679-753, floor policy :482-513, rank 8,
// 256 codewords).  Host and device: the kernels in k_lrq.hip call these per
// row, and the CPU tests compile the same header.
#pragma once
#include <stdint.h>

#include "dfloat.hpp"
#include "softdouble.hpp"

#define LRQ_RANKM 8   // compiled for rank 8 (the pinned ANQI_RANKM)

#if defined(__HIPCC__)
#define LRQ_SOFT __host__ __device__ __noinline__   // rare path: kept out of line
#else
#define LRQ_SOFT static
#endif

struct Coef8 {   // a row's coefficients, passed by value (registers)
    float v[LRQ_RANKM];
};

// std::upper_bound over 256 sorted codewords: the first index whose
// codeword is above v (256 when none is).  Eight halving steps settle the
// count among the first 255 codewords; the last one is checked after them.
FF_HD uint32_t lrq_upper_bound256(const float* cbk, float v) {
    uint32_t lo = 0;
#pragma unroll
    for (int sh = 7; sh >= 0; sh--) {
        const uint32_t step = 1u << sh;
        if (!(v < cbk[lo + step - 1])) lo += step;
    }
    if (!(v < cbk[lo])) lo++;   // only when lo == 255
    return lo;
}

// Soft binary64 coefficient: float(sum_k fma((double)r_k - mu_k, G_mk, c)).
LRQ_SOFT float lrq_coeff_soft(const float* rk, const float* mu, const float* gm, uint32_t nk) {
    uint64_t c = 0;
    for (uint32_t k = 0; k < nk; k++)
        c = sd_fma(sd_sub(sd_from_float(rk[k]), sd_from_float(mu[k])), sd_from_float(gm[k]), c);
    return sd_to_float(c);
}

// Soft binary64 floor code of rank k: upper_bound(cb_k, float(r_k - recon)) - 1.
LRQ_SOFT uint32_t lrq_code_soft(float rkk, Coef8 c, const float* basis, uint32_t nk, uint32_t k, float muk,
                                const float* cbk) {
    uint64_t recon = sd_from_float(muk);
    for (uint32_t m = 0; m < LRQ_RANKM; m++)
        recon = sd_fma(sd_from_float(c.v[m]), sd_from_float(basis[(size_t)m * nk + k]), recon);
    const uint32_t a = lrq_upper_bound256(cbk, sd_to_float(sd_sub(sd_from_float(rkk), recon)));
    return a == 0u ? 0u : a - 1u;
}

// Soft binary64 threshold of rank k, slack 1: float(max(0, mu_k + sum_m c_m G_mk + cw)),
// or, for the inner-product metric (clamp false), float(mu_k + sum_m c_m G_mk + cw).
LRQ_SOFT float lrq_thr_soft(Coef8 c, const float* basis, uint32_t nk, uint32_t k, float muk, float cw,
                            bool clamp = true) {
    uint64_t s = sd_from_float(muk);
    for (uint32_t m = 0; m < LRQ_RANKM; m++)
        s = sd_fma(sd_from_float(c.v[m]), sd_from_float(basis[(size_t)m * nk + k]), s);
    s = sd_add(s, sd_from_float(cw));
    return clamp && sd_lt_zero_or_zero(s) ? 0.0f : sd_to_float(s);
}

// Coefficients of row rk.  Binary64 error of the reference's sum: one
// rounding of each difference and of each fma, at most 2^-53 (nk + 1) sum|terms|.
// Each term (r_k - mu_k) G_mk = dh G + dl G (r - mu = dh + dl exactly) is
// split into the float-float's part p = fl(dh G) and a small remainder
// u = q + fl(dl G) (q = dh G - p exactly) of at most about 2^-23 |p|; the
// remainders add up in one float whose roundings (each at most 2^-24 of
// |the result|, and 2^-24 |fl(dl G)| for the product) join err, and that
// float joins the float-float once at the end.
struct LrqCoeffAcc {
    FFSum acc[LRQ_RANKM];
    float sm[LRQ_RANKM], se[LRQ_RANKM];
    int ok;
};

FF_HD void lrq_coeff_begin(LrqCoeffAcc& A) {
#pragma unroll
    for (int m = 0; m < LRQ_RANKM; m++) {
        A.acc[m] = ff_zero();
        A.sm[m] = 0.0f;
        A.se[m] = 0.0f;
    }
    A.ok = 1;
}

// The terms of rank position k (cutoff r, mean u).
FF_HD void lrq_coeff_term(LrqCoeffAcc& A, float r, float u, const float* basis, uint32_t nk, uint32_t k) {
    float dh, dl;
    ff_two_sum(r, -u, dh, dl);   // r - mu exactly
#pragma unroll
    for (int m = 0; m < LRQ_RANKM; m++) {
        const float g = basis[(size_t)m * nk + k];
        const float p = dh * g;
        const float q = fmaf(dh, g, -p);
        const float t = dl * g;
        // p in [2^-60, 2^100] (or 0) keeps q and t normal or exact
        A.ok &= (p == 0.0f || (fabsf(p) >= 0x1p-60f && fabsf(p) <= 0x1p100f)) &&
                (t == 0.0f ? (dl == 0.0f || g == 0.0f) : fabsf(t) >= 0x1p-100f);
        const float v = q + t;
        A.sm[m] += v;
        A.se[m] += fabsf(t) + fabsf(v) + fabsf(A.sm[m]);
        ff_add(A.acc[m], p);
    }
}

// Certified coefficients; uncertified ones recomputed from the row rk.
FF_HD Coef8 lrq_coeff_finish(LrqCoeffAcc& A, const float* rk, uint32_t nk, const float* mu, const float* basis,
                             uint32_t* fallbacks) {
    Coef8 c;
#pragma unroll
    for (int m = 0; m < LRQ_RANKM; m++) {
        A.acc[m].err += A.se[m];
        ff_add(A.acc[m], A.sm[m]);
        A.acc[m].ok &= A.ok;
        const float dref = A.acc[m].mag * ((float)(nk + 1) * 0x1.03p-53f);
        if (ff_round_certain(A.acc[m], dref)) {
            c.v[m] = A.acc[m].hi;
        } else {
            c.v[m] = lrq_coeff_soft(rk, mu, basis + (size_t)m * nk, nk);
            ++*fallbacks;
        }
    }
    return c;
}

FF_HD Coef8 lrq_coeffs(const float* rk, uint32_t nk, const float* mu, const float* basis, uint32_t* fallbacks) {
    LrqCoeffAcc A;
    lrq_coeff_begin(A);
    for (uint32_t k = 0; k < nk; k++) lrq_coeff_term(A, rk[k], mu[k], basis, nk, k);
    return lrq_coeff_finish(A, rk, nk, mu, basis, fallbacks);
}

// Code of rank k given the row's (float) coefficients.  Binary64 error of
// the reference's residual: rank + 1 roundings, each at most 2^-53
// (|r_k| + |mu_k| + sum_m |c_m G_mk|).  Only the codeword interval matters:
// when every float the residual can round to lands past the same codewords,
// the code is settled without knowing which float it is.
FF_HD uint32_t lrq_code(float rkk, Coef8 c, const float* basis, uint32_t nk, uint32_t k, float muk,
                        const float* cbk, uint32_t* fallbacks) {
    FFShort y = ffs_start(rkk);
    ffs_add(y, -muk);
#pragma unroll
    for (uint32_t m = 0; m < LRQ_RANKM; m++) ffs_add_prod(y, -c.v[m], basis[(size_t)m * nk + k]);
    const FFSum x = ffs_finish(y, LRQ_RANKM + 1, 2 * LRQ_RANKM + 1);
    const float dref = x.mag * ((float)(LRQ_RANKM + 1) * 0x1.03p-53f);
    float lo_f, hi_f;
    if (ff_float_interval(x, dref, lo_f, hi_f)) {
        const uint32_t a = lrq_upper_bound256(cbk, lo_f);
        if (a == 256u || !(cbk[a] <= hi_f)) return a == 0u ? 0u : a - 1u;
    }
    ++*fallbacks;
    return lrq_code_soft(rkk, c, basis, nk, k, muk, cbk);
}

// Threshold of rank k, slack 1.  Binary64 error: rank + 1 roundings, each at
// most 2^-53 (|mu_k| + sum_m |c_m G_mk| + |cw|).  clamp: the L2 metric's
// max(0, .); the inner-product metric keeps the sign (rknn_query.cpp:751).
FF_HD float lrq_thr(Coef8 c, const float* basis, uint32_t nk, uint32_t k, float muk, float cw,
                    uint32_t* fallbacks, bool clamp = true) {
    FFShort y = ffs_start(muk);
#pragma unroll
    for (uint32_t m = 0; m < LRQ_RANKM; m++) ffs_add_prod(y, c.v[m], basis[(size_t)m * nk + k]);
    ffs_add(y, cw);
    const FFSum x = ffs_finish(y, LRQ_RANKM + 1, 2 * LRQ_RANKM + 1);
    const float dref = x.mag * ((float)(LRQ_RANKM + 1) * 0x1.03p-53f);
    if (ff_round_certain(x, dref)) return !clamp || x.hi > 0.0f ? x.hi : 0.0f;   // clamped: a negative gives +0
    ++*fallbacks;
    return lrq_thr_soft(c, basis, nk, k, muk, cw, clamp);
}
