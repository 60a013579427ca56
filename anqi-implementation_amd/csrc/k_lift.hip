// k_lift.hip -- the binary64 lift of the base rows (src/rknn_query.cpp:2203-2236)
// on the device, with the integer soft binary64 of softdouble.hpp (no FP64
// instructions).  One thread per row; each thread runs its row's operations
// in the reference's order, so every per-row double is the host's value bit
// for bit.  The two reductions over rows that feed back into the rows (the
// sum of squared norms, in row order, and the maximum lifted norm) are done
// between the passes: the sum on the host, the maximum by an atomic maximum
// on the bit patterns (all values >= +0, where the bit order is the numeric
// order).
#include "common.hpp"
#include "kernels.hpp"
#include "softdouble.hpp"

// Pass 1: s_i = fma(o_ij, o_ij, s) over j in order, from s = +0.
__global__ void __launch_bounds__(256) k_lift_norm2(const float* __restrict__ x, uint64_t n, uint32_t d,
                                                    uint64_t* __restrict__ s_out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* o = x + i * d;
    uint64_t s = 0;
    for (uint32_t j = 0; j < d; j++) {
        const uint64_t v = sd_from_float(o[j]);
        s = sd_fma(v, v, s);
    }
    s_out[i] = s;
}

// Pass 2: v_j = o_j * scale, h_j = float(v_j), nr2 = fma(v_j, v_j, nr2);
// h_d = float(0.5 * fma(rk_i, scale^2, -nr2)) (L2) or float(rk_i * scale^2)
// (inner product, ip != 0); max of fma(h_d, h_d, nr2).
__global__ void __launch_bounds__(256) k_lift_rows(const float* __restrict__ x, uint64_t n, uint32_t d,
                                                   const float* __restrict__ rk, uint64_t scale, uint64_t sc2,
                                                   int ip, float* __restrict__ lifted,
                                                   unsigned long long* __restrict__ maxbits) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* o = x + i * d;
    float* h = lifted + i * (d + 2);
    uint64_t nr2 = 0;
    for (uint32_t j = 0; j < d; j++) {
        const uint64_t v = sd_mul(sd_from_float(o[j]), scale);
        h[j] = sd_to_float(v);
        nr2 = sd_fma(v, v, nr2);
    }
    float hd;
    if (ip) {
        hd = sd_to_float(sd_mul(sd_from_float(rk[i]), sc2));
    } else {
        const uint64_t t = sd_fma(sd_from_float(rk[i]), sc2, nr2 ^ (1ull << 63));
        hd = sd_to_float(sd_mul(t, 0x3fe0000000000000ull));   // 0.5 * t
    }
    h[d] = hd;
    const uint64_t hdd = sd_from_float(hd);
    const uint64_t hn2 = sd_fma(hdd, hdd, nr2);
    atomicMax(maxbits, (unsigned long long)hn2);
}

// Pass 3: hn2 = fma(h_j, h_j, hn2) over the d + 1 floats of the row;
// h_{d+1} = float(sqrt(max(0, M2 - hn2))).
__global__ void __launch_bounds__(256) k_lift_pad(float* __restrict__ lifted, uint64_t n, uint32_t d,
                                                  uint64_t m2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float* h = lifted + i * (d + 2);
    uint64_t hn2 = 0;
    for (uint32_t j = 0; j <= d; j++) {
        const uint64_t v = sd_from_float(h[j]);
        hn2 = sd_fma(v, v, hn2);
    }
    uint64_t pad = sd_sub(m2, hn2);
    if ((pad >> 63) && (pad & 0x7fffffffffffffffull)) pad = 0;        // pad < 0 -> 0
    h[d + 1] = sd_to_float(sd_sqrt(pad));
}

static unsigned lift_grid(uint64_t n) { return (unsigned)((n + 255) / 256); }

hipError_t launch_lift_norm2(const float* x, uint64_t n, uint32_t d, uint64_t* s_out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lift_norm2, dim3(lift_grid(n)), dim3(256), 0, s, x, n, d, s_out);
    return hipGetLastError();
}

hipError_t launch_lift_rows(const float* x, uint64_t n, uint32_t d, const float* rk, uint64_t scale, uint64_t sc2,
                            int ip, float* lifted, uint64_t* maxbits, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lift_rows, dim3(lift_grid(n)), dim3(256), 0, s, x, n, d, rk, scale, sc2, ip, lifted,
                       (unsigned long long*)maxbits);
    return hipGetLastError();
}

hipError_t launch_lift_pad(float* lifted, uint64_t n, uint32_t d, uint64_t m2, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lift_pad, dim3(lift_grid(n)), dim3(256), 0, s, lifted, n, d, m2);
    return hipGetLastError();
}

// AKNN keys -> verifier cutoffs and lift horizon (tools/build_aknn_source.cpp
// writes key = max(0, d2) for L2 into the .rec, :119-137; the lift's graph
// radius is the row's K-th cutoff, src/rknn_query.cpp:2203-2236): cut =
// max(0, keys) elementwise (NaN stays NaN, as std::max(0.0f, x) would keep
// 0 only for x < 0), horizon[i] = cut[i][K-1].
// For the "ip" metric (ipm) the .rec key is the distance minus 1
// (build_aknn_source.cpp:125-126), not clamped.
__global__ void __launch_bounds__(256) k_aknn_cutoffs(const float* __restrict__ keys, uint64_t m, uint32_t K,
                                                      int ipm, float* __restrict__ cut, float* __restrict__ horizon) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    const float v = keys[t];
    const float c = ipm ? v - 1.0f : (v < 0.0f ? 0.0f : v);
    cut[t] = c;
    if (horizon && t % K == K - 1) horizon[t / K] = c;
}

hipError_t launch_aknn_cutoffs(const float* keys, uint64_t n, uint32_t K, int ipm, float* cut, float* horizon,
                               hipStream_t s) {
    const uint64_t m = n * K;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_aknn_cutoffs, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, keys, m, K, ipm, cut,
                       horizon);
    return hipGetLastError();
}
