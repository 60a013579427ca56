// k_exact.hip -- exact self-excluded top-K (tools/dataset_gt.cpp:206-245) and
// exact reverse-kNN membership (dataset_gt.cpp:283-313) on the GPU.
//
// Keys follow dataset_gt: key = bn_i + bn_j - 2 <x_i, x_j> with the FP32
// contract of include/distance.h; the K smallest (key, id) pairs are kept
// (the strict-< max-heap over ascending j of the reference selects exactly
// these).
//
// Tiling: a block owns 64 rows; 16 groups of 16 lanes each hold 4 rows in
// registers (lane l: elements l, l+16, ...).  Column tiles of 64 rows are
// staged in LDS and every (row, column) pair of the tile is evaluated.  Pairs
// beating the row's current K-th pair go to a per-row LDS buffer (at most 64
// per tile, so it cannot overflow); after each tile each wavefront sorts its
// rows' buffers and merges them into the rows' sorted lists.
//
// The bf16 form (mfma != 0) evaluates the tile with v_mfma_f32_32x32x16_bf16
// on bf16 copies of the rows and keeps a pair for exact re-scoring when its
// approximate key lies below the row threshold plus a bound on the bf16 and
// accumulation error (|x||y| * margin_scale), so the output equals the FP32
// form.
#include "common.hpp"
#include "kernels.hpp"

#define EX_R 64        // rows per block
#define EX_C 64        // columns per tile
#define EX_RPG 4       // rows per 16-lane group

// ipm: the inner-product key -<x_i, x_j> (dataset_gt.cpp:225) instead of d2.
__global__ void __launch_bounds__(256) k_exact_topk(const float* __restrict__ x, const float* __restrict__ nrm,
                                                    uint32_t n, uint32_t d, uint32_t K, uint32_t row0,
                                                    uint32_t rows, int ipm, uint32_t* __restrict__ out_ids,
                                                    float* __restrict__ out_keys) {
    extern __shared__ __align__(16) char lds[];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, l = lane & 15;
    const int grp = tid >> 4;                      // 0..15
    const uint32_t rb = row0 + blockIdx.x * EX_R;  // first row of the block
    const uint32_t rend = min(row0 + rows, rb + EX_R);
    // LDS layout
    float* tile = (float*)lds;                                      // EX_C x d
    float* tnrm = tile + (size_t)EX_C * d;                          // EX_C
    float* ld = tnrm + EX_C;                                        // EX_R x K
    uint32_t* li = (uint32_t*)(ld + (size_t)EX_R * K);              // EX_R x K
    uint64_t* buf = (uint64_t*)(li + (size_t)EX_R * K);             // EX_R x 64
    uint32_t* bcnt = (uint32_t*)(buf + (size_t)EX_R * 64);          // EX_R
    uint32_t* lcnt = bcnt + EX_R;                                   // EX_R
    uint64_t* thr = (uint64_t*)(lcnt + EX_R + 2);                   // EX_R (8-aligned below)
    thr = (uint64_t*)(((uintptr_t)thr + 7) & ~(uintptr_t)7);
    float* sd_ = (float*)(thr + EX_R);                              // 4 waves x K
    uint32_t* si_ = (uint32_t*)(sd_ + 4 * (size_t)K);
    uint64_t* cs = (uint64_t*)(((uintptr_t)(si_ + 4 * (size_t)K) + 7) & ~(uintptr_t)7);  // 4 x 64

    for (int r = tid; r < EX_R; r += 256) { bcnt[r] = 0; lcnt[r] = 0; thr[r] = ~0ull; }
    // rows of this group in registers
    QRow QR[EX_RPG];
    float nr[EX_RPG];
    uint32_t rid[EX_RPG];
#pragma unroll
    for (int q = 0; q < EX_RPG; q++) {
        uint32_t r = rb + grp * EX_RPG + q;
        rid[q] = r;
        if (r < rend) { qrow_load(QR[q], x + (size_t)r * d, (int)d, l); nr[q] = nrm[r]; }
        else { for (int t = 0; t < QMAX_CHUNKS; t++) QR[q].v[t] = 0.0f; QR[q].tail = 0.0f; nr[q] = 0.0f; }
    }
    __syncthreads();
    for (uint32_t c0 = 0; c0 < n; c0 += EX_C) {
        uint32_t cn = min((uint32_t)EX_C, n - c0);
        // 16 independent loads in flight per thread before the LDS stores
        const uint32_t tot = cn * d;
        for (uint32_t b0 = 0; b0 < tot; b0 += 256 * 16) {
            float r[16];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                uint32_t e = b0 + (uint32_t)u * 256 + (uint32_t)tid;
                r[u] = e < tot ? x[(size_t)c0 * d + e] : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < 16; u++) {
                uint32_t e = b0 + (uint32_t)u * 256 + (uint32_t)tid;
                if (e < tot) tile[e] = r[u];
            }
        }
        for (uint32_t e = tid; e < cn; e += 256) tnrm[e] = nrm[c0 + e];
        __syncthreads();
        uint64_t th[EX_RPG];
#pragma unroll
        for (int q = 0; q < EX_RPG; q++) th[q] = thr[grp * EX_RPG + q];
        for (uint32_t c = 0; c < cn; c++) {
            const float* y = tile + (size_t)c * d;
            float nc = tnrm[c];
            uint32_t cid = c0 + c;
#pragma unroll
            for (int q = 0; q < EX_RPG; q++) {
                float ip = qrow_ip(QR[q], y, (int)d, l);
                float key = ipm ? -ip : d2_combine(ip, nr[q], nc);
                if (l == 0 && rid[q] < rend && cid != rid[q]) {
                    uint64_t pk = pkey(key, cid);
                    if (pk < th[q]) {
                        int r = grp * EX_RPG + q;
                        uint32_t p = atomicAdd(&bcnt[r], 1u);
                        buf[(size_t)r * 64 + p] = pk;
                    }
                }
            }
        }
        __syncthreads();
        // merge buffers: wavefront w owns rows w*16 .. w*16+15
        for (int rr = 0; rr < 16; rr++) {
            int r = w * 16 + rr;
            uint32_t bc = bcnt[r];
            if (bc == 0) continue;
            uint64_t key = lane < (int)bc ? buf[(size_t)r * 64 + lane] : ~0ull;
            key = wave_sort64(key);
            cs[w * 64 + lane] = key;
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
            float* rd = ld + (size_t)r * K;
            uint32_t* ri = li + (size_t)r * K;
            float* od = sd_ + (size_t)w * K;
            uint32_t* oi = si_ + (size_t)w * K;
            uint32_t size = lcnt[r];
            const uint64_t* csw = cs + w * 64;
            for (uint32_t i = lane; i < size; i += 64) {
                uint64_t bk = pkey(rd[i], ri[i]);
                int lo = 0, hi = (int)bc;
                while (lo < hi) { int mid = (lo + hi) >> 1; if (csw[mid] < bk) lo = mid + 1; else hi = mid; }
                uint32_t pos = i + (uint32_t)lo;
                if (pos < K) { od[pos] = rd[i]; oi[pos] = ri[i]; }
            }
            if (lane < (int)bc) {
                uint32_t lo = 0, hi = size;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (pkey(rd[mid], ri[mid]) < key) lo = mid + 1; else hi = mid;
                }
                uint32_t pos = (uint32_t)lane + lo;
                if (pos < K) { od[pos] = pkey_d(key); oi[pos] = pkey_id(key); }
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
            uint32_t ns = min(K, size + bc);
            for (uint32_t i = lane; i < ns; i += 64) { rd[i] = od[i]; ri[i] = oi[i]; }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) {
                lcnt[r] = ns;
                bcnt[r] = 0;
                if (ns >= K) thr[r] = pkey(rd[K - 1], ri[K - 1]);
            }
        }
        __syncthreads();
    }
    for (int r = w * 16; r < w * 16 + 16; r++) {
        uint32_t row = rb + r;
        if (row >= rend) continue;
        for (uint32_t i = lane; i < K; i += 64) {
            out_ids[(size_t)(row - row0) * K + i] = i < lcnt[r] ? li[(size_t)r * K + i] : U32_NONE;
            out_keys[(size_t)(row - row0) * K + i] = i < lcnt[r] ? ld[(size_t)r * K + i] : F32_BIG;
        }
    }
}

static size_t exact_lds(uint32_t d, uint32_t K) {
    return sizeof(float) * ((size_t)EX_C * d + EX_C) + (sizeof(float) + sizeof(uint32_t)) * (size_t)EX_R * K +
           sizeof(uint64_t) * EX_R * 64 + sizeof(uint32_t) * (2 * EX_R + 2) + 8 + sizeof(uint64_t) * EX_R +
           (sizeof(float) + sizeof(uint32_t)) * 4 * (size_t)K + 8 + sizeof(uint64_t) * 4 * 64 + 64;
}

hipError_t launch_exact_topk(const float* x, const float* nrm, uint32_t n, uint32_t d,
                             uint32_t K, uint32_t row0, uint32_t rows, int ipm, uint32_t* ids,
                             float* keys, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    size_t lds = exact_lds(d, K);
    if (hipError_t e = set_lds_limit((const void*)k_exact_topk, lds); e != hipSuccess) return e;
    hipLaunchKernelGGL(k_exact_topk, dim3((rows + EX_R - 1) / EX_R), dim3(256), lds, s, x, nrm, n, d, K, row0,
                       rows, ipm, ids, keys);
    return hipGetLastError();
}

// ---- reverse-kNN membership ------------------------------------------------------------
#define RK_COLS 4096
// ipm: the inner-product key -<q, x_j> (dataset_gt.cpp:292) instead of d2.
template <bool FILL>
__global__ void __launch_bounds__(256) k_rknn(const float* __restrict__ base, const float* __restrict__ bnorm,
                                              uint32_t n, const float* __restrict__ q, uint32_t d, int ipm,
                                              const float* __restrict__ rk2, uint32_t* __restrict__ counts,
                                              const uint32_t* __restrict__ off, uint32_t* __restrict__ ids,
                                              float* __restrict__ keys) {
    const uint32_t qi = blockIdx.y;
    const int tid = threadIdx.x, l = tid & 15, grp = tid >> 4;
    QRow Q;
    qrow_load(Q, q + (size_t)qi * d, (int)d, l);
    float qn = qrow_ip(Q, q + (size_t)qi * d, (int)d, l);
    uint32_t c0 = blockIdx.x * RK_COLS, c1 = min(n, c0 + RK_COLS);
    uint32_t local = 0;
    for (uint32_t j = c0 + grp; j < c1; j += 16) {
        float ip = qrow_ip(Q, base + (size_t)j * d, (int)d, l);
        float key = ipm ? -ip : fmaf(-2.0f, ip, qn + bnorm[j]);
        if (l == 0 && key <= rk2[j]) {
            if (FILL) {
                uint32_t p = atomicAdd(&counts[qi], 1u);
                ids[off[qi] + p] = j;
                keys[off[qi] + p] = key;
            } else {
                local++;
            }
        }
    }
    if (!FILL && l == 0 && local) atomicAdd(&counts[qi], local);
}

hipError_t launch_rknn_count(const float* base, const float* bnorm, uint32_t n,
                             const float* q, uint32_t nq, uint32_t d, int ipm, const float* rk2,
                             uint32_t* counts, hipStream_t s) {
    if (nq == 0 || n == 0) return hipSuccess;
    dim3 grid((n + RK_COLS - 1) / RK_COLS, nq);
    hipLaunchKernelGGL((k_rknn<false>), grid, dim3(256), 0, s, base, bnorm, n, q, d, ipm, rk2, counts,
                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (float*)nullptr);
    return hipGetLastError();
}

hipError_t launch_rknn_fill(const float* base, const float* bnorm, uint32_t n,
                            const float* q, uint32_t nq, uint32_t d, int ipm, const float* rk2,
                            const uint32_t* off, uint32_t* ids, float* keys, hipStream_t s) {
    if (nq == 0 || n == 0) return hipSuccess;
    uint32_t* cursor = nullptr;
    hipError_t e = hipMallocAsync((void**)&cursor, sizeof(uint32_t) * nq, s);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(cursor, 0, sizeof(uint32_t) * nq, s);
    if (e == hipSuccess) {
        dim3 grid((n + RK_COLS - 1) / RK_COLS, nq);
        hipLaunchKernelGGL((k_rknn<true>), grid, dim3(256), 0, s, base, bnorm, n, q, d, ipm, rk2, cursor, off, ids,
                           keys);
        e = hipGetLastError();
    }
    const hipError_t ef = hipFreeAsync(cursor, s);
    return e != hipSuccess ? e : ef;
}

__global__ void k_to_bf16(const float* __restrict__ x, uint64_t m, uint16_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    uint32_t u = __float_as_uint(x[i]);
    uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);   // round to nearest even
    out[i] = (uint16_t)(r >> 16);
}

hipError_t launch_to_bf16(const float* x, uint64_t m, uint16_t* out, hipStream_t s) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, x, m, out);
    return hipGetLastError();
}
