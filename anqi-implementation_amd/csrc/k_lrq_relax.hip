// k_lrq_relax.hip -- differentiable radius-reconstruction objective of the
// LRQ verifier fit, forward and hand-written backward.
//
// The reference trainer (tools/train_final_m8_u8.py:81-156) alternates a
// PCA basis with 1-D Lloyd codebooks on the residuals; its objective is the
// radius reconstruction r ~ mu + c G + cb[code] with c = (r - mu) G^T
// (:92-94) and nearest-codeword residual decoding (:146-156).  Here that
// objective is relaxed so that it can be optimised by gradient descent: the
// codeword choice becomes a softmax over the 256 codewords of each rank,
//
//   y = r - mu,  c_m = sum_k y_k G_mk,  p_k = sum_m c_m G_mk,  e_k = y_k - p_k,
//   w_kj = softmax_j(-(e_k - cb_kj)^2 / tau),  q_k = sum_j w_kj cb_kj,
//   z_k = e_k - q_k,   loss = mean_ik z_ik^2,
//
// and the exact floor decode (k_lrq.hip) stays the arbiter of the fitted
// basis and codebook.  Backward, per row (g = 2 z * scale):
//   dq/de = 2/tau Var_w(cb)              -> ge_k = g_k (1 - dq_k/de_k)
//   dL/dcb_kj = -g_k w_kj (1 + 2/tau (cb_kj - q_k)(e_k - cb_kj))
//   gp = -ge,  gc_m = sum_k gp_k G_mk,  dL/dG_mk = c_m gp_k + gc_m y_k
// Floor mode (mode 1) relaxes the verifier's own decode instead: the largest
// codeword <= e becomes, for sorted codewords and sigma = logistic,
//   q = cb_0 + sum_{j>=1} sigma((e - cb_j)/beta) (cb_j - cb_{j-1})
// (exactly the floor decode as beta -> 0), with
//   dq/de = sum_j s_j (1 - s_j) D_j / beta,  s_j = sigma((e - cb_j)/beta),
//   dq/dcb_m = [m=0] + [m>=1] s_m - [m<=J-2] s_{m+1} - [m>=1] s_m (1 - s_m) D_m / beta
// (tau carries beta).
// Pass A (one block per row) produces c, e, q, softmax max/sum, ge, gc and
// the row loss; pass B (block per rank and row chunk, one thread per
// codeword) accumulates dL/dcb; pass C (block per row chunk, one thread per
// rank) accumulates dL/dG.  FP32 throughout.
#include "common.hpp"
#include "kernels.hpp"

#define RX_T 128          // threads of a pass-A block (ranks k < nk <= 128)
#define RX_MMAX 16        // rank-M <= 16

__device__ __forceinline__ float rx_block_sum(float v, float* red) {
    // RX_T = 2 wavefronts
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    return red[0] + red[1];
}

__global__ void __launch_bounds__(RX_T) k_relax_rows(const float* __restrict__ cut, uint32_t n, uint32_t nk,
                                                     uint32_t rankm, const float* __restrict__ mu,
                                                     const float* __restrict__ G, const float* __restrict__ cb,
                                                     uint32_t ncode, float inv_tau, float gscale, int mode,
                                                     float* __restrict__ c_out, float* __restrict__ e_out,
                                                     float* __restrict__ q_out, float* __restrict__ mx_out,
                                                     float* __restrict__ sm_out, float* __restrict__ ge_out,
                                                     float* __restrict__ gc_out, float* __restrict__ loss_rows) {
    __shared__ float red[2];
    const uint32_t k = threadIdx.x;
    const bool on = k < nk;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const float y = on ? cut[(size_t)i * nk + k] - mu[k] : 0.0f;
        float c[RX_MMAX];
        for (uint32_t m = 0; m < rankm; m++) c[m] = rx_block_sum(on ? y * G[(size_t)m * nk + k] : 0.0f, red);
        float p = 0.0f;
        for (uint32_t m = 0; m < rankm; m++) p = fmaf(c[m], on ? G[(size_t)m * nk + k] : 0.0f, p);
        const float e = y - p;
        float ge = 0.0f, z = 0.0f;
        if (on && mode == 1) {
            const float* cbk = cb + (size_t)k * ncode;
            float qsum = 0.0f, dq = 0.0f;
            for (uint32_t j = 1; j < ncode; j++) {
                const float D = cbk[j] - cbk[j - 1];
                const float sj = 1.0f / (1.0f + __expf(-(e - cbk[j]) * inv_tau));
                qsum = fmaf(sj, D, qsum);
                dq = fmaf(sj * (1.0f - sj), D, dq);
            }
            z = (e - cbk[0]) - qsum;
            const float g = 2.0f * z * gscale;
            ge = g * (1.0f - dq * inv_tau);
            const size_t o = (size_t)i * nk + k;
            e_out[o] = e;
            q_out[o] = z;
            ge_out[o] = ge;
        } else if (on) {
            const float* cbk = cb + (size_t)k * ncode;
            float mx = -__builtin_huge_valf();
            for (uint32_t j = 0; j < ncode; j++) {
                const float t = e - cbk[j];
                mx = fmaxf(mx, -t * t * inv_tau);
            }
            // mean and variance of u = cb - e under w, the variance about the
            // mean in a separate pass (no cancellation when e lies far from
            // every codeword)
            float s0 = 0.0f, s1 = 0.0f;
            for (uint32_t j = 0; j < ncode; j++) {
                const float u = cbk[j] - e;
                const float w = __expf(-u * u * inv_tau - mx);
                s0 += w;
                s1 = fmaf(w, u, s1);
            }
            const float mu1 = s1 / s0;               // q - e
            float s2 = 0.0f;
            for (uint32_t j = 0; j < ncode; j++) {
                const float u = cbk[j] - e;
                const float w = __expf(-u * u * inv_tau - mx);
                const float v = u - mu1;
                s2 = fmaf(w * v, v, s2);
            }
            const float var = s2 / s0;
            z = -mu1;                                 // e - q without cancellation
            const float g = 2.0f * z * gscale;
            ge = g * (1.0f - 2.0f * inv_tau * var);
            const size_t o = (size_t)i * nk + k;
            e_out[o] = e;
            q_out[o] = z;
            mx_out[o] = mx;
            sm_out[o] = s0;
            ge_out[o] = ge;
        }
        const float lrow = rx_block_sum(z * z, red);
        for (uint32_t m = 0; m < rankm; m++) {
            const float gc = rx_block_sum(on ? -ge * G[(size_t)m * nk + k] : 0.0f, red);
            if (k == 0) {
                c_out[(size_t)i * rankm + m] = c[m];
                gc_out[(size_t)i * rankm + m] = gc;
            }
        }
        if (k == 0) loss_rows[i] = lrow;
    }
}

#define RX_CHUNK 512
__global__ void k_relax_gcb(const float* __restrict__ e, const float* __restrict__ q, const float* __restrict__ mx,
                            const float* __restrict__ sm, const float* __restrict__ cb, uint32_t n, uint32_t nk,
                            uint32_t ncode, float inv_tau, float gscale, int mode, float* __restrict__ gcb) {
    const uint32_t k = blockIdx.x, j = threadIdx.x;
    if (j >= ncode) return;
    const float v = cb[(size_t)k * ncode + j];
    const uint32_t i0 = blockIdx.y * RX_CHUNK, i1 = min(n, i0 + RX_CHUNK);
    float acc = 0.0f;
    if (mode == 1) {
        const float Dm = j >= 1 ? v - cb[(size_t)k * ncode + j - 1] : 0.0f;
        const float vn = j + 1 < ncode ? cb[(size_t)k * ncode + j + 1] : 0.0f;
        for (uint32_t i = i0; i < i1; i++) {
            const size_t o = (size_t)i * nk + k;
            const float ee = e[o];
            const float g = 2.0f * q[o] * gscale;     // q holds z
            float dqd;
            if (j == 0) {
                dqd = 1.0f - (ncode > 1 ? 1.0f / (1.0f + __expf(-(ee - vn) * inv_tau)) : 0.0f);
            } else {
                const float sm_ = 1.0f / (1.0f + __expf(-(ee - v) * inv_tau));
                const float sn = j + 1 < ncode ? 1.0f / (1.0f + __expf(-(ee - vn) * inv_tau)) : 0.0f;
                dqd = sm_ - sn - sm_ * (1.0f - sm_) * Dm * inv_tau;
            }
            acc = fmaf(-g, dqd, acc);
        }
        atomicAdd(&gcb[(size_t)k * ncode + j], acc);
        return;
    }
    for (uint32_t i = i0; i < i1; i++) {
        const size_t o = (size_t)i * nk + k;
        const float ee = e[o], zz = q[o];           // q holds z = e - q
        const float t = ee - v;
        const float w = __expf(-t * t * inv_tau - mx[o]) / sm[o];
        const float g = 2.0f * zz * gscale;
        acc = fmaf(-g * w, 1.0f + 2.0f * inv_tau * (zz - t) * t, acc);   // cb - q = z - t
    }
    atomicAdd(&gcb[(size_t)k * ncode + j], acc);
}

__global__ void k_relax_gG(const float* __restrict__ cut, const float* __restrict__ mu,
                           const float* __restrict__ c, const float* __restrict__ gc, const float* __restrict__ ge,
                           uint32_t n, uint32_t nk, uint32_t rankm, float* __restrict__ gG) {
    const uint32_t k = threadIdx.x;
    if (k >= nk) return;
    const uint32_t i0 = blockIdx.x * RX_CHUNK, i1 = min(n, i0 + RX_CHUNK);
    float acc[RX_MMAX];
    for (uint32_t m = 0; m < rankm; m++) acc[m] = 0.0f;
    const float muk = mu[k];
    for (uint32_t i = i0; i < i1; i++) {
        const float y = cut[(size_t)i * nk + k] - muk;
        const float gp = -ge[(size_t)i * nk + k];
        for (uint32_t m = 0; m < rankm; m++)
            acc[m] = fmaf(c[(size_t)i * rankm + m], gp, fmaf(gc[(size_t)i * rankm + m], y, acc[m]));
    }
    for (uint32_t m = 0; m < rankm; m++) atomicAdd(&gG[(size_t)m * nk + k], acc[m]);
}

hipError_t launch_relax(const RelaxParams& P, hipStream_t s) {
    if (P.nk > RX_T || P.rankm > RX_MMAX || P.ncode > 1024 || P.n == 0) return hipErrorInvalidValue;
    const uint32_t grid = std::min<uint32_t>(P.n, 65536u);
    hipLaunchKernelGGL(k_relax_rows, dim3(grid), dim3(RX_T), 0, s, P.cut, P.n, P.nk, P.rankm, P.mu, P.G, P.cb,
                       P.ncode, P.inv_tau, P.gscale, P.mode, P.c, P.e, P.q, P.mx, P.sm, P.ge, P.gc, P.loss_rows);
    if (P.gcb) {
        dim3 g2(P.nk, (P.n + RX_CHUNK - 1) / RX_CHUNK);
        hipLaunchKernelGGL(k_relax_gcb, g2, dim3((P.ncode + 63) / 64 * 64), 0, s, P.e, P.q, P.mx, P.sm, P.cb, P.n,
                           P.nk, P.ncode, P.inv_tau, P.gscale, P.mode, P.gcb);
    }
    if (P.gG) {
        hipLaunchKernelGGL(k_relax_gG, dim3((P.n + RX_CHUNK - 1) / RX_CHUNK), dim3(RX_T), 0, s, P.cut, P.mu, P.c,
                           P.gc, P.ge, P.n, P.nk, P.rankm, P.gG);
    }
    return hipGetLastError();
}
