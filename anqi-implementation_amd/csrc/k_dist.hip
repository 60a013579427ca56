// k_dist.hip -- row norms and pairwise d2 under the reference FP32 contract
// (include/distance.h:143-161 and include/index.h:172-176).
#include "common.hpp"
#include "kernels.hpp"

// One 16-lane group per row.
__global__ void __launch_bounds__(256) k_norms(const float* __restrict__ x, uint64_t n, uint32_t d,
                                               float* __restrict__ out) {
    uint64_t row = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    int l = threadIdx.x & 15;
    if (row >= n) return;
    const float* r = x + row * d;
    float v = g16_ip(r, r, (int)d, l);
    if (l == 0) out[row] = v;
}

__global__ void __launch_bounds__(256) k_pair_d2(const float* __restrict__ x, const float* __restrict__ nrm,
                                                 uint32_t d, const uint32_t* __restrict__ a,
                                                 const uint32_t* __restrict__ b, uint64_t m,
                                                 float* __restrict__ out) {
    uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    int l = threadIdx.x & 15;
    if (t >= m) return;
    uint32_t i = a[t], j = b[t];
    float ip = g16_ip(x + (size_t)i * d, x + (size_t)j * d, (int)d, l);
    if (l == 0) out[t] = d2_combine(ip, nrm[i], nrm[j]);
}

hipError_t launch_norms(const float* x, uint64_t n, uint32_t d, float* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_norms, dim3((unsigned)((n * 16 + 255) / 256)), dim3(256), 0, s, x, n, d, out);
    return hipGetLastError();
}

hipError_t launch_pair_d2(const float* x, const float* nrm, uint32_t d, const uint32_t* a,
                          const uint32_t* b, uint64_t m, float* out, hipStream_t s) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_d2, dim3((unsigned)((m * 16 + 255) / 256)), dim3(256), 0, s, x, nrm, d, a, b,
                       m, out);
    return hipGetLastError();
}
