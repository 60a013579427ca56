// Host build of the verifier's certified encode/decode (anqi-implementation_amd/
// csrc/lrq_cert.hpp), the per-row functions the GPU kernels call, for the CPU
// tests (tests/test_lrq_cert_host.py compiles it with -ffp-contract=off).
#include <stddef.h>
#include <stdint.h>

#include "lrq_cert.hpp"

extern "C" {

uint32_t host_upper_bound256(const float* cbk, float v) { return lrq_upper_bound256(cbk, v); }

// Row-major codes (n x nk); returns the elements recomputed in soft binary64.
uint64_t host_lrq_encode(const float* cut, uint64_t n, uint32_t nk, const float* mu, const float* basis,
                         const float* cb, float* coeff, uint8_t* codes) {
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t fb = 0;
        const float* rk = cut + i * nk;
        const Coef8 c = lrq_coeffs(rk, nk, mu, basis, &fb);
        for (uint32_t k = 0; k < nk; k++)
            codes[i * nk + k] = (uint8_t)lrq_code(rk[k], c, basis, nk, k, mu[k], cb + (size_t)k * 256, &fb);
        for (uint32_t m = 0; m < LRQ_RANKM; m++) coeff[i * LRQ_RANKM + m] = c.v[m];
        total += fb;
    }
    return total;
}

uint64_t host_lrq_decode(const float* coeff, const uint8_t* codes, uint64_t n, uint32_t nk, const float* mu,
                         const float* basis, const float* cb, uint32_t k_one_based, float* thr) {
    uint64_t total = 0;
    const uint32_t k = k_one_based - 1;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t fb = 0;
        Coef8 c;
        for (uint32_t m = 0; m < LRQ_RANKM; m++) c.v[m] = coeff[i * LRQ_RANKM + m];
        thr[i] = lrq_thr(c, basis, nk, k, mu[k], cb[(size_t)k * 256 + codes[i * nk + k]], &fb);
        total += fb;
    }
    return total;
}

// The float-float estimates taken without their certificates: the planted
// mistake the tests must catch (uncertified roundings differ from binary64).
void host_lrq_uncertified_coeffs(const float* cut, uint64_t n, uint32_t nk, const float* mu, const float* basis,
                                 float* coeff) {
    for (uint64_t i = 0; i < n; i++) {
        const float* rk = cut + i * nk;
        for (uint32_t m = 0; m < LRQ_RANKM; m++) {
            FFSum x = ff_zero();
            for (uint32_t k = 0; k < nk; k++) {
                float dh, dl;
                ff_two_sum(rk[k], -mu[k], dh, dl);
                ff_add_prod(x, dh, basis[(size_t)m * nk + k]);
                ff_add(x, dl * basis[(size_t)m * nk + k]);
            }
            coeff[i * LRQ_RANKM + m] = x.hi;
        }
    }
}

}  // extern "C"
