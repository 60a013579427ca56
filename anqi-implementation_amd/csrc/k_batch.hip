// k_batch.hip -- result compaction for captured query batches: the accepted
// ids of every query of a batch, in candidate (discovery) order, as one CSR
// (offsets + ids), so a batch returns only its answers.  The batch size is
// read from device memory, so a captured graph serves batches of any size up
// to its capacity.
#include "common.hpp"
#include "kernels.hpp"

// One block: exclusive scan of pred_cnt[0..nq) into off[0..nq], nq <= max_nq.
__global__ void __launch_bounds__(1024) k_batch_scan(const uint32_t* __restrict__ pred, const uint32_t* nq_dev,
                                                     uint32_t* __restrict__ off) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const uint32_t nq = *nq_dev;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < nq; b += 1024) {
        const uint32_t i = b + (uint32_t)tid;
        const uint32_t v = i < nq ? pred[i] : 0u;
        // wave inclusive scan
        uint32_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint32_t before = carry;
        for (int j = 0; j < w; j++) before += wsum[j];
        if (i < nq) off[i] = before + x - v;
        __syncthreads();
        if (tid == 0) {
            uint32_t t = 0;
            for (int j = 0; j < 16; j++) t += wsum[j];
            carry += t;
        }
        __syncthreads();
    }
    if (tid == 0) off[nq] = carry;
}

// One wavefront per query.
__global__ void __launch_bounds__(64) k_batch_compact(const uint32_t* __restrict__ cand, const uint8_t* __restrict__ acc,
                                                      const uint32_t* __restrict__ cnt, const uint32_t* __restrict__ off,
                                                      const uint32_t* nq_dev, uint32_t out_cap,
                                                      uint32_t* __restrict__ ids) {
    const uint32_t q = blockIdx.x;
    if (q >= *nq_dev) return;
    const int lane = __lane_id();
    const uint32_t c = min(cnt[q], out_cap);
    uint32_t pos = off[q];
    for (uint32_t b = 0; b < c; b += 64) {
        const uint32_t t = b + (uint32_t)lane;
        const int ok = t < c && acc[(size_t)q * out_cap + t];
        const uint64_t m = ballot(ok);
        if (ok) ids[pos + lanes_below(m)] = cand[(size_t)q * out_cap + t];
        pos += (uint32_t)popc64(m);
    }
}

hipError_t launch_batch_scan(const uint32_t* pred_cnt, const uint32_t* nq_dev, uint32_t max_nq, uint32_t* off,
                             hipStream_t s) {
    (void)max_nq;
    hipLaunchKernelGGL(k_batch_scan, dim3(1), dim3(1024), 0, s, pred_cnt, nq_dev, off);
    return hipGetLastError();
}

hipError_t launch_batch_compact(const uint32_t* cand_ids, const uint8_t* accept, const uint32_t* cand_cnt,
                                const uint32_t* off, const uint32_t* nq_dev, uint32_t max_nq, uint32_t out_cap,
                                uint32_t* ids, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_compact, dim3(max_nq), dim3(64), 0, s, cand_ids, accept, cand_cnt, off, nq_dev,
                       out_cap, ids);
    return hipGetLastError();
}

// ---- candidate compaction (the drop-in tool's answers) --------------------------
// Every query's candidates (first min(cnt, out_cap) of its row) and their
// accept bytes as one CSR, so the host copies what the search found instead
// of nq x out_cap rows.  One block scans the row lengths; one wavefront per
// query copies.
__global__ void __launch_bounds__(1024) k_cand_scan(const uint32_t* __restrict__ cnt, uint32_t nq, uint32_t out_cap,
                                                    uint32_t* __restrict__ off) {
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t carry;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (uint32_t b = 0; b < nq; b += 1024) {
        const uint32_t i = b + (uint32_t)tid;
        const uint32_t v = i < nq ? min(cnt[i], out_cap) : 0u;
        uint32_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        uint32_t before = carry;
        for (int j = 0; j < w; j++) before += wsum[j];
        if (i < nq) off[i] = before + x - v;
        __syncthreads();
        if (tid == 0) {
            uint32_t t = 0;
            for (int j = 0; j < 16; j++) t += wsum[j];
            carry += t;
        }
        __syncthreads();
    }
    if (tid == 0) off[nq] = carry;
}

__global__ void __launch_bounds__(64) k_cand_copy(const uint32_t* __restrict__ cand, const uint8_t* __restrict__ acc,
                                                  const uint32_t* __restrict__ off, uint32_t out_cap,
                                                  uint32_t* __restrict__ ids, uint8_t* __restrict__ acc_out) {
    const uint32_t q = blockIdx.x;
    const uint32_t c = off[q + 1] - off[q];
    for (uint32_t t = __lane_id(); t < c; t += 64) {
        ids[off[q] + t] = cand[(size_t)q * out_cap + t];
        acc_out[off[q] + t] = acc[(size_t)q * out_cap + t];
    }
}

hipError_t launch_cand_scan(const uint32_t* cnt, uint32_t nq, uint32_t out_cap, uint32_t* off, hipStream_t s) {
    hipLaunchKernelGGL(k_cand_scan, dim3(1), dim3(1024), 0, s, cnt, nq, out_cap, off);
    return hipGetLastError();
}

hipError_t launch_cand_copy(const uint32_t* cand, const uint8_t* acc, const uint32_t* off, uint32_t nq,
                            uint32_t out_cap, uint32_t* ids, uint8_t* acc_out, hipStream_t s) {
    if (nq == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cand_copy, dim3(nq), dim3(64), 0, s, cand, acc, off, out_cap, ids, acc_out);
    return hipGetLastError();
}
