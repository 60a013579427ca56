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

// The butterfly partners are taken with DPP row rotations (row_ror 8, 4, 2,
// 1 within each 16-lane row) instead of xor shuffles through the LDS
// crossbar.  After the step with distance 8, lanes i and i^8 hold equal
// values, after distance 4 also i^4, and so on; so lane (i + s) mod 16 holds
// the same value as lane i ^ s at every step, and each add sees the operands
// of the xor tree (IEEE addition is commutative): the bits are those of
// _mm512_reduce_add_ps.  Whole 16-lane groups are active or inactive together
// at every call site.
__device__ __forceinline__ float dpp_ror16(float v, int s) {
    const int x = __float_as_int(v);
    int r;
    switch (s) {
        case 8: r = __builtin_amdgcn_mov_dpp(x, 0x128, 0xf, 0xf, false); break;
        case 4: r = __builtin_amdgcn_mov_dpp(x, 0x124, 0xf, 0xf, false); break;
        case 2: r = __builtin_amdgcn_mov_dpp(x, 0x122, 0xf, 0xf, false); break;
        default: r = __builtin_amdgcn_mov_dpp(x, 0x121, 0xf, 0xf, false); break;
    }
    return __int_as_float(r);
}

__device__ __forceinline__ float tree16(float v) {
    v = v + dpp_ror16(v, 8);
    v = v + dpp_ror16(v, 4);
    v = v + dpp_ror16(v, 2);
    v = v + dpp_ror16(v, 1);
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

// qrow_load / qrow_ip / qrow_l2 for rows of d = 16 MC + r (r < 16) with MC
// known at compile time: every element of the row is requested before the
// first one is used, so a row costs one memory round trip (with a run-time
// chunk count each chunk's load sits in its own branch next to its FMA and
// the loads go out one after another).  Same FMA order, reduction tree and
// tail term as the run-time forms, hence the same bits.  MC = 0: the
// run-time forms.
template <int MC>
__device__ __forceinline__ void qrow_load_mc(QRow& q, const float* __restrict__ x, int d, int l) {
    if constexpr (MC == 0) {
        qrow_load(q, x, d, l);
    } else {
        const int r = d - 16 * MC;
#pragma unroll
        for (int t = 0; t < QMAX_CHUNKS; t++) q.v[t] = (t < MC) ? x[16 * t + l] : 0.0f;
        q.tail = (l < r) ? x[16 * MC + l] : 0.0f;
    }
}

template <int MC>
__device__ __forceinline__ float qrow_ip_mc(const QRow& q, const float* __restrict__ y, int d, int l) {
    if constexpr (MC == 0) {
        return qrow_ip(q, y, d, l);
    } else {
        const int r = d - 16 * MC;
        float yv[MC];
#pragma unroll
        for (int t = 0; t < MC; t++) yv[t] = y[16 * t + l];
        const float yt = (l < r) ? y[16 * MC + l] : 0.0f;
        float acc = 0.0f;
#pragma unroll
        for (int t = 0; t < MC; t++) acc = fmaf(q.v[t], yv[t], acc);
        float res = tree16(acc);
        if (r) {
            float p = (l < r) ? q.tail * yt : 0.0f;
            res += tree16(p);
        }
        return res;
    }
}

template <int MC>
__device__ __forceinline__ float qrow_l2_mc(const QRow& q, const float* __restrict__ y, int d, int l) {
    if constexpr (MC == 0) {
        return qrow_l2(q, y, d, l);
    } else {
        const int r = d - 16 * MC;
        float yv[MC];
#pragma unroll
        for (int t = 0; t < MC; t++) yv[t] = y[16 * t + l];
        const float yt = (l < r) ? y[16 * MC + l] : 0.0f;
        float acc = 0.0f;
#pragma unroll
        for (int t = 0; t < MC; t++) {
            const float a = q.v[t] - yv[t];
            acc = fmaf(a, a, acc);
        }
        float res = tree16(acc);
        if (r) {
            const float a = (l < r) ? q.tail - yt : 0.0f;
            res += tree16(a * a);
        }
        return res;
    }
}

// Batched row distances: for rows ids[s], s < k, group g (16 lanes) takes
// s = g, g+4, ..., g + 4 (RB - 1) of each 4 RB-row pass and issues all RB
// rows' loads before reducing any, so the fetches overlap; the last pass
// halves RB while the rows left fit, so a short batch issues no loads or
// arithmetic for absent rows.  Same per-lane FMA order and reduction tree as
// qrow_l2 / qrow_ip, hence the same bits.
// MC = number of full 16-float chunks of a row of X (dim = 16 MC + r, r < 16).
//   MODE 0: out[s] = ||q - x||^2
//   MODE 1: out[s] = d2_combine(<q, x>, nq, nrm[id])   (nndindex::d2)
//   MODE 2: out[s] = ||q - x||^2 over the dim (= 16 MC) coordinates of X's rows
//           + nrm[id] (the rest of the sum, already reduced: split lifted rows)
//   MODE 3: out[s] = -<q, x>   (the inner-product key, dataset_gt.cpp:225)
template <int MC, int MODE, int RB>
__device__ __forceinline__ void rows_pass(const QRow& q, const float* __restrict__ X, uint32_t dim,
                                          const float* __restrict__ nrm, float nq, const uint32_t* ids, int k,
                                          int l, int g, float* out, int s0) {
    const int r = (int)dim - 16 * MC;
    float x[RB][MC];
    float xt[RB], nb[RB];
#pragma unroll
    for (int j = 0; j < RB; j++) {
        const int s = s0 + g + 4 * j;
        if (s < k) {
            const uint32_t id = ids[s];
            const float* row = X + (size_t)id * dim;
#pragma unroll
            for (int t = 0; t < MC; t++) x[j][t] = row[16 * t + l];
            xt[j] = (l < r) ? row[16 * MC + l] : 0.0f;
            nb[j] = (MODE == 1 || MODE == 2) ? nrm[id] : 0.0f;
        } else {
#pragma unroll
            for (int t = 0; t < MC; t++) x[j][t] = 0.0f;
            xt[j] = 0.0f;
            nb[j] = 0.0f;
        }
    }
#pragma unroll
    for (int j = 0; j < RB; j++) {
        float acc = 0.0f;
        float res;
        if (MODE != 1 && MODE != 3) {
#pragma unroll
            for (int t = 0; t < MC; t++) {
                float a = q.v[t] - x[j][t];
                acc = fmaf(a, a, acc);
            }
            res = tree16(acc);
            if (MODE == 0 && r) {
                float a = q.tail - xt[j];
                res += tree16(a * a);
            }
            if (MODE == 2) res += nb[j];
        } else {
#pragma unroll
            for (int t = 0; t < MC; t++) acc = fmaf(q.v[t], x[j][t], acc);
            res = tree16(acc);
            if (r) res += tree16(q.tail * xt[j]);
            res = MODE == 3 ? -res : fmaf(-2.0f, res, nq + nb[j]);
        }
        const int s = s0 + g + 4 * j;
        if (l == 0 && s < k) out[s] = res;
    }
}

// the last pass: 0 < k - s0 <= 4 RB rows
template <int MC, int MODE, int RB>
__device__ __forceinline__ void rows_tail(const QRow& q, const float* __restrict__ X, uint32_t dim,
                                          const float* __restrict__ nrm, float nq, const uint32_t* ids, int k,
                                          int l, int g, float* out, int s0) {
    if constexpr (RB > 1) {
        if (k - s0 <= 2 * RB) {
            rows_tail<MC, MODE, RB / 2>(q, X, dim, nrm, nq, ids, k, l, g, out, s0);
            return;
        }
    }
    rows_pass<MC, MODE, RB>(q, X, dim, nrm, nq, ids, k, l, g, out, s0);
}

template <int MC, int MODE, int RB = 4>
__device__ __forceinline__ void rows_batch(const QRow& q, const float* __restrict__ X, uint32_t dim,
                                           const float* __restrict__ nrm, float nq,
                                           const uint32_t* ids, int k, int l, int g, float* out) {
    int s0 = 0;
    for (; k - s0 > 4 * RB; s0 += 4 * RB) rows_pass<MC, MODE, RB>(q, X, dim, nrm, nq, ids, k, l, g, out, s0);
    if (s0 < k) rows_tail<MC, MODE, RB>(q, X, dim, nrm, nq, ids, k, l, g, out, s0);
}

// nndindex's "ip" distance 1 - <x, y> (include/distance.h:170-175), a scalar
// loop in the reference that GCC 11 at the reference's flags vectorises: the
// products of the first nv = 16 floor(d / 16) + (8 if d mod 16 >= 8) elements
// are rounded and added to the sum one at a time in element order, the last
// d mod 8 run as fused multiply-adds.  One lane per pair (the sum is
// sequential); rows 16-byte aligned when d % 4 == 0.
__device__ __forceinline__ float ipdist_seq(const float* __restrict__ x, const float* __restrict__ y, uint32_t d) {
    uint32_t nv = 16u * (d / 16u);
    if (d - nv >= 8u) nv += 8u;
    float p = 0.0f;
    if ((d & 3u) == 0u) {
        for (uint32_t i = 0; i < nv; i += 4) {
            const float4 a = *reinterpret_cast<const float4*>(x + i);
            const float4 b = *reinterpret_cast<const float4*>(y + i);
            p = p + a.x * b.x;
            p = p + a.y * b.y;
            p = p + a.z * b.z;
            p = p + a.w * b.w;
        }
    } else {
        for (uint32_t i = 0; i < nv; i++) p = p + x[i] * y[i];
    }
    for (uint32_t i = nv; i < d; i++) p = fmaf(x[i], y[i], p);
    return 1.0f - p;
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

// ---- register-resident sorted key lists -------------------------------------
// A list of up to 64 NCH (dist, id) keys, sorted ascending: lane l holds
// positions 64 j + l, j < NCH (~0 past the end), plus one bit per chunk (x)
// for "expanded".  Merging kc candidate keys needs no sort and no LDS search:
// for each candidate (scalar loop) one ballot per chunk counts the list keys
// below it and every list lane counts the candidates below its key; the rank
// among the candidates is one more ballot.  Keys then move to their new
// positions through one LDS write and one read per chunk.  Single wavefront.
#define RK_FLAG 0x80000000ull   // expanded bit, kept in bit 31 of the id while in LDS

template <int NCH>
struct RKeys {
    uint64_t k[NCH];
    uint32_t x;   // bit j: position 64 j + lane expanded
};

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int src) {
    const uint32_t lo = __builtin_amdgcn_readlane((int)(uint32_t)v, src);
    const uint32_t hi = __builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

// LDS ordering between lanes of one wavefront.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
}

template <int NCH>
__device__ __forceinline__ void rk_clear(RKeys<NCH>& B) {
#pragma unroll
    for (int j = 0; j < NCH; j++) B.k[j] = ~0ull;
    B.x = 0;
}

// key at a (wave-uniform) position
template <int NCH>
__device__ __forceinline__ uint64_t rk_get(const RKeys<NCH>& B, uint32_t pos) {
    uint64_t v = ~0ull;
#pragma unroll
    for (int j = 0; j < NCH; j++)
        if ((uint32_t)j == (pos >> 6)) v = readlane64(B.k[j], (int)(pos & 63));
    return v;
}

// Merges the lanes' candidate keys (~0 = none; none equal to a list key)
// into the list of `size` entries, capacity cap <= 64 NCH: candidates that
// cannot enter a full list are dropped first.  kb: LDS exchange area of cap
// entries.  Returns the number of candidates that entered the merge.
template <int NCH>
__device__ __forceinline__ uint32_t rk_merge(RKeys<NCH>& B, uint64_t key, uint32_t& size, uint32_t cap,
                                             uint64_t* kb) {
    const int lane = __lane_id();
    if (size == cap && !(key < rk_get(B, cap - 1))) key = ~0ull;
    const uint64_t km = ballot(key != ~0ull);
    if (km == 0) return 0;
    uint32_t sc[NCH];
#pragma unroll
    for (int j = 0; j < NCH; j++) sc[j] = 0;
    uint32_t mypos = U32_NONE;
    for (uint64_t mm = km; mm; mm &= mm - 1) {
        const int src = __ffsll((long long)mm) - 1;
        const uint64_t cv = readlane64(key, src);
        uint32_t pos = (uint32_t)popc64(ballot(key < cv));
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            pos += (uint32_t)popc64(ballot(B.k[j] < cv));
            sc[j] += (cv < B.k[j]) ? 1u : 0u;
        }
        if (lane == src) mypos = pos;
    }
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        const uint32_t np = 64u * j + (uint32_t)lane + sc[j];
        if (B.k[j] != ~0ull && np < cap) kb[np] = B.k[j] | (((B.x >> j) & 1u) ? RK_FLAG : 0ull);
    }
    if (mypos < cap) kb[mypos] = key;
    const uint32_t kc = (uint32_t)popc64(km);
    size = min(cap, size + kc);
    wave_lds_sync();
    B.x = 0;
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        const uint32_t pos = 64u * j + (uint32_t)lane;
        uint64_t v = ~0ull;
        if (pos < size) {
            v = kb[pos];
            if (v & RK_FLAG) B.x |= 1u << j;
            v &= ~RK_FLAG;
        }
        B.k[j] = v;
    }
    wave_lds_sync();
    return kc;
}

// rk_merge with a 32-bit exchange area (cap words, half the LDS): the keys
// move in two passes, high words then low words.  The positions are those of
// rk_merge, so the list is the same.  Between the passes a lane's register
// holds the new high word and the old low word it still has to move.
template <int NCH>
__device__ __forceinline__ uint32_t rk_merge_x32(RKeys<NCH>& B, uint64_t key, uint32_t& size, uint32_t cap,
                                                 uint32_t* kb) {
    const int lane = __lane_id();
    if (size == cap && !(key < rk_get(B, cap - 1))) key = ~0ull;
    const uint64_t km = ballot(key != ~0ull);
    if (km == 0) return 0;
    uint32_t sc[NCH];
#pragma unroll
    for (int j = 0; j < NCH; j++) sc[j] = 0;
    uint32_t mypos = U32_NONE;
    for (uint64_t mm = km; mm; mm &= mm - 1) {
        const int src = __ffsll((long long)mm) - 1;
        const uint64_t cv = readlane64(key, src);
        uint32_t pos = (uint32_t)popc64(ballot(key < cv));
#pragma unroll
        for (int j = 0; j < NCH; j++) {
            pos += (uint32_t)popc64(ballot(B.k[j] < cv));
            sc[j] += (cv < B.k[j]) ? 1u : 0u;
        }
        if (lane == src) mypos = pos;
    }
    const uint32_t kc = (uint32_t)popc64(km);
    const uint32_t nsize = min(cap, size + kc);
    // the old entries (ov bit j: present) with their expanded flags (bit 31
    // of the id word)
    uint32_t ov = 0;
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        if (B.k[j] != ~0ull) {
            ov |= 1u << j;
            if ((B.x >> j) & 1u) B.k[j] |= RK_FLAG;
        }
    }
    // pass 1: high words
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        const uint32_t np = 64u * j + (uint32_t)lane + sc[j];
        if (((ov >> j) & 1u) && np < cap) kb[np] = (uint32_t)(B.k[j] >> 32);
    }
    if (mypos < cap) kb[mypos] = (uint32_t)(key >> 32);
    wave_lds_sync();
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        const uint32_t pos = 64u * j + (uint32_t)lane;
        const uint32_t hi = pos < nsize ? kb[pos] : ~0u;
        B.k[j] = ((uint64_t)hi << 32) | (uint32_t)B.k[j];   // new high word, old low word
    }
    wave_lds_sync();
    // pass 2: low words
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        const uint32_t np = 64u * j + (uint32_t)lane + sc[j];
        if (((ov >> j) & 1u) && np < cap) kb[np] = (uint32_t)B.k[j];
    }
    if (mypos < cap) kb[mypos] = (uint32_t)key;
    wave_lds_sync();
    B.x = 0;
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        const uint32_t pos = 64u * j + (uint32_t)lane;
        uint64_t v = ~0ull;
        if (pos < nsize) {
            v = (B.k[j] & 0xFFFFFFFF00000000ull) | kb[pos];
            if (v & RK_FLAG) B.x |= 1u << j;
            v &= ~RK_FLAG;
        }
        B.k[j] = v;
    }
    size = nsize;
    wave_lds_sync();
    return kc;
}

// Id of the first unexpanded entry (U32_NONE if none); mark: flag it expanded.
template <int NCH>
__device__ __forceinline__ uint32_t rk_take(RKeys<NCH>& B, bool mark) {
    const int lane = __lane_id();
    int cj = -1, cl = 0;
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        if (cj < 0) {
            const uint64_t m = ballot(B.k[j] != ~0ull && !((B.x >> j) & 1u));
            if (m) { cj = j; cl = __ffsll((long long)m) - 1; }
        }
    }
    if (cj < 0) return U32_NONE;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < NCH; j++) {
        if (j == cj) {
            c = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)B.k[j], cl);
            if (mark && lane == cl) B.x |= 1u << j;
        }
    }
    return c;
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
// mask + 1 slots (a multiple of 4) in buckets of four (16 bytes), probed
// bucket by bucket: one 16-byte read answers "present" or finds a free slot,
// so a lookup costs one LDS round trip unless its bucket is full.  Slots are
// never freed, so a key lives in the first bucket of its probe sequence that
// had room when it was inserted; every bucket before it is still full.  Lanes
// inserting different keys into one bucket race through the CAS; a lane that
// loses moves to the next free slot.  (One call never holds a key twice.)
__device__ __forceinline__ uint32_t hmix(uint32_t id) {
    uint32_t h = id * 0x9E3779B1u;
    return h ^ (h >> 16);
}
__device__ __forceinline__ uint32_t hslot(uint32_t id, uint32_t mask) { return hmix(id) & mask; }
// home bucket of id among nb buckets (any nb)
__device__ __forceinline__ uint32_t hbucket(uint32_t id, uint32_t nb) {
    return (uint32_t)(((uint64_t)hmix(id) * nb) >> 32);
}
// Returns 1 if id was inserted (first visit), 0 if present, -1 if table full.
__device__ __forceinline__ int hset_insert(uint32_t* tab, uint32_t mask, uint32_t id) {
    const uint32_t key = id + 1u;
    const uint32_t nb = (mask + 1u) >> 2;
    uint32_t b = hbucket(id, nb);
    for (uint32_t probe = 0; probe < nb; probe++) {
        uint32_t* bk = tab + 4u * b;
        const uint4 v = *(const uint4*)bk;
        if (v.x == key || v.y == key || v.z == key || v.w == key) return 0;
        uint32_t e = v.x == 0u ? 0u : v.y == 0u ? 1u : v.z == 0u ? 2u : v.w == 0u ? 3u : 4u;
        for (; e < 4u; e++) {
            const uint32_t old = atomicCAS(&bk[e], 0u, key);
            if (old == 0u) return 1;
            if (old == key) return 0;
        }
        b = (b + 1u == nb) ? 0u : b + 1u;
    }
    return -1;
}
// ---- compact LDS id set: 16-bit entries, ids < 2^20 -----------------------------
// 4 << lb entries in 1 << lb buckets of four (8 bytes), lb = 8..11.  h =
// mix20(id) is a bijection of the 20-bit ids; its top lb bits pick the home
// bucket and an entry stores (probe distance << (20 - lb)) | (low 20 - lb bits
// of h): with the bucket it sits in that gives back h, hence the id, so the
// set is exact.  Probe distances 0 .. 2^(lb - 4) - 2; the all-ones entry
// 0xFFFF (the next distance) marks an empty slot.  Probing as in hset_insert:
// a key lives in the first bucket of its sequence that had room, so a bucket
// with an empty entry ends a lookup.  A bucket is updated with one 64-bit
// CAS; a lane whose CAS loses re-reads the bucket.
#define H16_EMPTY 0xFFFFu
#define H16_MAX_ID (1u << 20)
#define H24_MAX_ID (1u << 24)
#define H24_OVF_SLOTS 64u
#define H16_LB_GATHER 11u
#define H16_SLOTS (4u << H16_LB_GATHER)
// the AKNN gather's 20-bit table: 4,096 entries (8 KB), so more rows fit a CU
#define H16_LB_AKNN 10u
#define H16_SLOTS_AKNN (4u << H16_LB_AKNN)
__host__ __device__ __forceinline__ uint32_t mix20(uint32_t id) {
    uint32_t h = (id * 0x9E3779B1u) & 0xFFFFFu;
    h ^= h >> 10;
    return (h * 0x2C1B3C6Du) & 0xFFFFFu;
}
// Returns 1 if id was inserted (first visit), 0 if present, -1 if its probe
// sequence is full.
__device__ __forceinline__ int hset16_insert(uint16_t* tab, uint32_t lb, uint32_t id) {
    const uint32_t rb = 20u - lb, h = mix20(id);
    const uint32_t home = h >> rb, rem = h & ((1u << rb) - 1u);
    const uint32_t maxd = (1u << (16u - rb)) - 1u, bmask = (1u << lb) - 1u;
    for (uint32_t dist = 0; dist < maxd; dist++) {
        const uint32_t code = (dist << rb) | rem;
        unsigned long long* bk = (unsigned long long*)(tab + 4u * ((home + dist) & bmask));
        unsigned long long v = *bk;
        for (;;) {
            int fe = -1;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint32_t x = (uint32_t)(v >> (16 * e)) & 0xFFFFu;
                if (x == code) return 0;
                if (x == H16_EMPTY && fe < 0) fe = e;
            }
            if (fe < 0) break;                                   // full: next distance
            const unsigned long long nv =
                (v & ~(0xFFFFull << (16 * fe))) | ((unsigned long long)code << (16 * fe));
            const unsigned long long old = atomicCAS(bk, v, nv);
            if (old == v) return 1;
            v = old;
        }
    }
    return -1;
}

// The same 16-bit table for ids < 2^24, with a 32-bit overflow table.  h =
// mix24(id) (a bijection of the 24-bit ids); the top 11 bits pick the home
// bucket of 2048 and an entry stores (probe distance << 13) | (low 13 bits of
// h), probe distances 0..6 (0xFFFF is distance 7: empty).  A key whose seven
// buckets were full when it came is kept in ovf, H24_OVF_SLOTS 32-bit slots
// (id + 1, 0 = empty) probed linearly.  A bucket with an empty entry ends a
// lookup as before; seven full buckets send it on to ovf, which holds every
// such key.
__host__ __device__ __forceinline__ uint32_t mix24(uint32_t id) {
    uint32_t h = (id * 0x9E3779B1u) & 0xFFFFFFu;
    h ^= h >> 11;
    return (h * 0x2C1B3C6Du) & 0xFFFFFFu;
}
__device__ __forceinline__ int hset24_insert(uint16_t* tab, uint32_t* ovf, uint32_t id) {
    const uint32_t h = mix24(id);
    const uint32_t home = h >> 13, rem = h & 0x1FFFu;
    for (uint32_t dist = 0; dist < 7u; dist++) {
        const uint32_t code = (dist << 13) | rem;
        unsigned long long* bk = (unsigned long long*)(tab + 4u * ((home + dist) & 2047u));
        unsigned long long v = *bk;
        for (;;) {
            int fe = -1;
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const uint32_t x = (uint32_t)(v >> (16 * e)) & 0xFFFFu;
                if (x == code) return 0;
                if (x == H16_EMPTY && fe < 0) fe = e;
            }
            if (fe < 0) break;
            const unsigned long long nv =
                (v & ~(0xFFFFull << (16 * fe))) | ((unsigned long long)code << (16 * fe));
            const unsigned long long old = atomicCAS(bk, v, nv);
            if (old == v) return 1;
            v = old;
        }
    }
    const uint32_t key = id + 1u;
    uint32_t sl = hmix(id) & (H24_OVF_SLOTS - 1u);
    for (uint32_t probe = 0; probe < H24_OVF_SLOTS; probe++) {
        const uint32_t old = atomicCAS(&ovf[sl], 0u, key);
        if (old == 0u) return 1;
        if (old == key) return 0;
        sl = (sl + 1u) & (H24_OVF_SLOTS - 1u);
    }
    return -1;
}
// First key of an empty table (the search's entry point).
__device__ __forceinline__ void hset_first(uint32_t* tab, uint32_t mask, uint32_t id) {
    tab[4u * hbucket(id, (mask + 1u) >> 2)] = id + 1u;
}

#define HIP_CHECK_RET(expr)                                                   \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess) return anqi_fail_hip(ctx, _e, #expr, __LINE__); \
    } while (0)

// Raises a kernel's dynamic LDS limit before its launch (launchers return
// the error instead of launching).
static inline hipError_t set_lds_limit(const void* f, size_t bytes) {
    return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}
