// common.hpp -- device building blocks shared by the ANQI kernels (gfx950).
//
// Distances follow the reference's FP32 contract (include/distance.h:22-161
// compiled by GCC): 16 accumulation lanes with fused multiply-add over full
// 16-float chunks, GCC's _mm512_reduce_add_ps tree, and a separately reduced
// tail of plain products.  On a 64-wide wavefront a 16-lane group owns one
// pair: lane l accumulates elements l, l+16, ..., so the four butterfly
// shuffles (xor 8, 4, 2, 1) reproduce the reduction tree exactly (each add is
// commutative).  All code is built with -ffp-contract=off; fmaf() is spelled
// where the reference fuses.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "softdouble.hpp"

#define ANQI_WAVE 64
#define U32_NONE 0xFFFFFFFFu
#define F32_BIG 3.402823466e+38f

// ---- 16-lane group helpers ------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ int grp_lane() { return __lane_id() & 15; }
__device__ __forceinline__ int grp_id() { return __lane_id() >> 4; }

__device__ __forceinline__ float tree16(float v) {
    v = v + __shfl_xor(v, 8, 16);
    v = v + __shfl_xor(v, 4, 16);
    v = v + __shfl_xor(v, 2, 16);
    v = v + __shfl_xor(v, 1, 16);
    return v;
}

// <x, y> over d floats, computed by the 16 lanes of a group (l = lane & 15).
// Every lane of the group returns the same bits.
__device__ __forceinline__ float g16_ip(const float* __restrict__ x,
                                        const float* __restrict__ y, int d, int l) {
    int m = d >> 4, r = d & 15;
    float acc = 0.0f;
    int t = 0;
    for (; t + 4 <= m; t += 4) {
        float a0 = x[16 * t + l], b0 = y[16 * t + l];
        float a1 = x[16 * (t + 1) + l], b1 = y[16 * (t + 1) + l];
        float a2 = x[16 * (t + 2) + l], b2 = y[16 * (t + 2) + l];
        float a3 = x[16 * (t + 3) + l], b3 = y[16 * (t + 3) + l];
        acc = fmaf(a0, b0, acc);
        acc = fmaf(a1, b1, acc);
        acc = fmaf(a2, b2, acc);
        acc = fmaf(a3, b3, acc);
    }
    for (; t < m; t++) acc = fmaf(x[16 * t + l], y[16 * t + l], acc);
    float res = tree16(acc);
    if (r) {
        float p = (l < r) ? x[16 * m + l] * y[16 * m + l] : 0.0f;
        res += tree16(p);
    }
    return res;
}

// ||x - y||^2 (distance.h:58-86).
__device__ __forceinline__ float g16_l2(const float* __restrict__ x,
                                        const float* __restrict__ y, int d, int l) {
    int m = d >> 4, r = d & 15;
    float acc = 0.0f;
    int t = 0;
    for (; t + 4 <= m; t += 4) {
        float a0 = x[16 * t + l] - y[16 * t + l];
        float a1 = x[16 * (t + 1) + l] - y[16 * (t + 1) + l];
        float a2 = x[16 * (t + 2) + l] - y[16 * (t + 2) + l];
        float a3 = x[16 * (t + 3) + l] - y[16 * (t + 3) + l];
        acc = fmaf(a0, a0, acc);
        acc = fmaf(a1, a1, acc);
        acc = fmaf(a2, a2, acc);
        acc = fmaf(a3, a3, acc);
    }
    for (; t < m; t++) {
        float a = x[16 * t + l] - y[16 * t + l];
        acc = fmaf(a, a, acc);
    }
    float res = tree16(acc);
    if (r) {
        float a = (l < r) ? x[16 * m + l] - y[16 * m + l] : 0.0f;
        res += tree16(a * a);
    }
    return res;
}

// Register-resident query row: lane l of a group holds x[16t + l], t < 16
// (d <= 256) plus the tail element.
#define QMAX_CHUNKS 16
struct QRow {
    float v[QMAX_CHUNKS];
    float tail;
};

__device__ __forceinline__ void qrow_load(QRow& q, const float* __restrict__ x, int d, int l) {
    int m = d >> 4, r = d & 15;
#pragma unroll
    for (int t = 0; t < QMAX_CHUNKS; t++) q.v[t] = (t < m) ? x[16 * t + l] : 0.0f;
    q.tail = (l < r) ? x[16 * m + l] : 0.0f;
}

__device__ __forceinline__ float qrow_ip(const QRow& q, const float* __restrict__ y, int d, int l) {
    int m = d >> 4, r = d & 15;
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < QMAX_CHUNKS; t++)
        if (t < m) acc = fmaf(q.v[t], y[16 * t + l], acc);
    float res = tree16(acc);
    if (r) {
        float p = (l < r) ? q.tail * y[16 * m + l] : 0.0f;
        res += tree16(p);
    }
    return res;
}

__device__ __forceinline__ float qrow_l2(const QRow& q, const float* __restrict__ y, int d, int l) {
    int m = d >> 4, r = d & 15;
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < QMAX_CHUNKS; t++)
        if (t < m) {
            float a = q.v[t] - y[16 * t + l];
            acc = fmaf(a, a, acc);
        }
    float res = tree16(acc);
    if (r) {
        float a = (l < r) ? q.tail - y[16 * m + l] : 0.0f;
        res += tree16(a * a);
    }
    return res;
}

// Batched row distances: for rows ids[s], s < k, group g (16 lanes) takes
// s = g, g+4, g+8, g+12 of each 16-row pass and issues all four rows' loads
// before reducing any, so the fetches overlap.  Same per-lane FMA order and
// reduction tree as qrow_l2 / qrow_ip, hence the same bits.  MC = number of
// full 16-float chunks of a row of X (dim = 16 MC + r, r < 16).
//   MODE 0: out[s] = ||q - x||^2
//   MODE 1: out[s] = d2_combine(<q, x>, nq, nrm[id])   (nndindex::d2)
template <int MC, int MODE>
__device__ __forceinline__ void rows_batch(const QRow& q, const float* __restrict__ X, uint32_t dim,
                                           const float* __restrict__ nrm, float nq,
                                           const uint32_t* ids, int k, int l, int g, float* out) {
    const int r = (int)dim - 16 * MC;
    for (int s0 = 0; s0 < k; s0 += 16) {
        float x[4][MC];
        float xt[4], nb[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int s = s0 + g + 4 * j;
            if (s < k) {
                const uint32_t id = ids[s];
                const float* row = X + (size_t)id * dim;
#pragma unroll
                for (int t = 0; t < MC; t++) x[j][t] = row[16 * t + l];
                xt[j] = (l < r) ? row[16 * MC + l] : 0.0f;
                nb[j] = MODE == 1 ? nrm[id] : 0.0f;
            } else {
#pragma unroll
                for (int t = 0; t < MC; t++) x[j][t] = 0.0f;
                xt[j] = 0.0f;
                nb[j] = 0.0f;
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            float acc = 0.0f;
            float res;
            if (MODE == 0) {
#pragma unroll
                for (int t = 0; t < MC; t++) {
                    float a = q.v[t] - x[j][t];
                    acc = fmaf(a, a, acc);
                }
                res = tree16(acc);
                if (r) {
                    float a = q.tail - xt[j];
                    res += tree16(a * a);
                }
            } else {
#pragma unroll
                for (int t = 0; t < MC; t++) acc = fmaf(q.v[t], x[j][t], acc);
                res = tree16(acc);
                if (r) res += tree16(q.tail * xt[j]);
                res = fmaf(-2.0f, res, nq + nb[j]);
            }
            const int s = s0 + g + 4 * j;
            if (l == 0 && s < k) out[s] = res;
        }
    }
}

// nndindex::d2 (include/index.h:172-176): na + nb - 2 ip.
__device__ __forceinline__ float d2_combine(float ip, float na, float nb) {
    return fmaf(-2.0f, ip, na + nb);
}

// ---- ordering keys ----------------------------------------------------------
// Monotone map float -> uint32 (total order matching '<' on non-NaN floats,
// -0 and +0 adjacent).
__device__ __forceinline__ uint32_t fkey(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}
// (dist, id) lexicographic as one 64-bit key.
__device__ __forceinline__ uint64_t pkey(float d, uint32_t id) {
    return ((uint64_t)fkey(d) << 32) | id;
}
__device__ __forceinline__ float pkey_d(uint64_t k) { return fkey_inv((uint32_t)(k >> 32)); }
__device__ __forceinline__ uint32_t pkey_id(uint64_t k) { return (uint32_t)k; }

// ---- wave utilities ---------------------------------------------------------
__device__ __forceinline__ uint64_t ballot(int p) { return __ballot(p); }
__device__ __forceinline__ int popc64(uint64_t m) { return __popcll(m); }
// number of set bits of m strictly below this lane
__device__ __forceinline__ int lanes_below(uint64_t m) {
    return __popcll(m & ((1ull << __lane_id()) - 1ull));
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    uint32_t lo = __shfl((uint32_t)v, src), hi = __shfl((uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
    uint32_t lo = __shfl_xor((uint32_t)v, m), hi = __shfl_xor((uint32_t)(v >> 32), m);
    return ((uint64_t)hi << 32) | lo;
}

// Bitonic sort of one 64-bit key per lane across the wavefront (ascending).
__device__ __forceinline__ uint64_t wave_sort64(uint64_t v) {
    int lane = __lane_id();
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            uint64_t o = shfl_xor64(v, j);
            bool up = ((lane & k) == 0);
            bool lower = ((lane & j) == 0);
            uint64_t mn = v < o ? v : o, mx = v < o ? o : v;
            v = (lower == up) ? mn : mx;
        }
    }
    return v;
}

// ---- counter-based RNG ------------------------------------------------------
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z ^= z >> 30;
    z *= 0xBF58476D1CE4E5B9ull;
    z ^= z >> 27;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 31;
    return z;
}
__host__ __device__ __forceinline__ uint64_t rng_at(uint64_t seed, uint64_t a, uint64_t b) {
    return mix64(mix64(seed ^ 0x9E3779B97F4A7C15ull) + a * 0xD1B54A32D192ED03ull + b * 0x9E3779B97F4A7C15ull);
}

// ---- LDS open-addressing id set (key = id + 1, 0 = empty) --------------------
__device__ __forceinline__ uint32_t hslot(uint32_t id, uint32_t mask) {
    uint32_t h = id * 0x9E3779B1u;
    return (h ^ (h >> 16)) & mask;
}
// Returns 1 if id was inserted (first visit), 0 if present, -1 if table full.
__device__ __forceinline__ int hset_insert(uint32_t* tab, uint32_t mask, uint32_t id) {
    uint32_t key = id + 1u;
    uint32_t s = hslot(id, mask);
    for (uint32_t probe = 0; probe <= mask; probe++) {
        uint32_t old = atomicCAS(&tab[s], 0u, key);
        if (old == 0u) return 1;
        if (old == key) return 0;
        s = (s + 1u) & mask;
    }
    return -1;
}

#define HIP_CHECK_RET(expr)                                                   \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess) return anqi_fail_hip(ctx, _e, #expr, __LINE__); \
    } while (0)
