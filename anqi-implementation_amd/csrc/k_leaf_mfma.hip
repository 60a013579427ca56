// k_leaf_mfma.hip -- RP-forest leaf all-pairs (init_knn, include/index.h:
// 1433-1449) with the pair sweep on the matrix cores.
//
// Same result as k_leaf_knn (the FP32 sweep): for every row of the leaf, the
// leaf members whose exact (key, id) beats the row's current K-th pair are
// merged into its sorted top-K list.  Exact keys use the reference's FP32
// contract; the bf16 MFMA sweep only skips pairs that provably cannot beat
// that threshold (approx > K-th key + M_r, with M_r the per-row bound of
// k_exact_margin).  Survivors are re-scored exactly; several flushes per row
// and tree are allowed because a candidate rejected by a tighter threshold
// could not have stayed among the K best anyway.
//
// One block (4 wavefronts) per leaf.  A wavefront takes 32-member row blocks
// of the leaf; for each it sweeps the leaf's 32-member column tiles with
// v_mfma_f32_32x32x16_bf16 (operands gathered from a bf16 copy of the base
// through the tree permutation), keeps per-row survivor lists in its own LDS
// region and flushes them (exact re-score, dedup against the list, merge)
// when a list could overflow on the next tile, and at the end.
#include "common.hpp"
#include "kernels.hpp"

typedef __bf16 lm_bf16x8 __attribute__((ext_vector_type(8)));
typedef float lm_f32x16 __attribute__((ext_vector_type(16)));

#define LM_KS 8     // 16-wide k steps: d <= 128, zero padded
#define LM_SV 64    // survivor slots per row

__host__ __device__ inline size_t lm_r16(size_t b) { return (b + 15) & ~size_t(15); }

// per-wave LDS region
__host__ __device__ inline size_t lm_wave_bytes(uint32_t K) {
    return lm_r16(2ull * 32 * LM_SV) + lm_r16(4ull * 32) + lm_r16(4ull * 64) + lm_r16(4ull * 64) +
           4 * lm_r16(4ull * K) + lm_r16(8ull * 64);
}

#define LM_WSYNC()                                              \
    do {                                                        \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  \
        __builtin_amdgcn_wave_barrier();                        \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  \
    } while (0)

struct LmWave {
    uint16_t* sv;    // 32 x LM_SV member indices
    uint32_t* svc;   // 32 counts
    uint32_t* xid;   // 64
    float* xd;       // 64
    float* ld0;      // K
    uint32_t* li0;
    float* ld1;
    uint32_t* li1;
    uint64_t* cs;    // 64
};

__device__ inline LmWave lm_carve(char* base, uint32_t K) {
    LmWave s;
    size_t o = 0;
    auto take = [&](size_t b) { char* p = base + o; o += lm_r16(b); return p; };
    s.sv = (uint16_t*)take(2ull * 32 * LM_SV);
    s.svc = (uint32_t*)take(4ull * 32);
    s.xid = (uint32_t*)take(4ull * 64);
    s.xd = (float*)take(4ull * 64);
    s.ld0 = (float*)take(4ull * K);
    s.li0 = (uint32_t*)take(4ull * K);
    s.ld1 = (float*)take(4ull * K);
    s.li1 = (uint32_t*)take(4ull * K);
    s.cs = (uint64_t*)take(8ull * 64);
    return s;
}

// Exact re-score of row r's survivors, dedup against its list, merge.
template <int MC>
__device__ void lm_flush_row(const LeafParams& P, LmWave& W, const uint32_t* mem, int rl, uint32_t r) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    const uint32_t K = P.K, d = P.d;
    const int k = (int)W.svc[rl];
    if (lane < k) W.xid[lane] = mem[W.sv[rl * LM_SV + lane]];
    LM_WSYNC();
    QRow QR;
    qrow_load(QR, P.x + (size_t)r * d, (int)d, l);
    const float nr = P.nrm[r];
    rows_batch<MC, 1>(QR, P.x, d, P.nrm, nr, W.xid, k, l, g, W.xd);
    LM_WSYNC();
    uint32_t* rid = P.ids + (size_t)r * K;
    float* rkd = P.keys + (size_t)r * K;
    const uint32_t cnt = P.cnt[r];
    const uint64_t thr = cnt >= K ? pkey(rkd[K - 1], rid[K - 1]) : ~0ull;
    uint64_t key = ~0ull;
    if (lane < k) {
        uint64_t pk = pkey(W.xd[lane], W.xid[lane]);
        if (pk < thr) key = pk;
    }
    key = wave_sort64(key);
    const int nc = popc64(ballot(key != ~0ull));
    if (lane == 0) W.svc[rl] = 0;
    if (nc == 0) { LM_WSYNC(); return; }
    for (uint32_t t = lane; t < cnt; t += 64) { W.ld0[t] = rkd[t]; W.li0[t] = rid[t]; }
    LM_WSYNC();
    // drop candidates already present (identical pair -> identical bits)
    int keep = 0;
    if (lane < nc) {
        uint32_t lo = 0, hi = cnt;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (pkey(W.ld0[mid], W.li0[mid]) < key) lo = mid + 1; else hi = mid;
        }
        keep = !(lo < cnt && W.li0[lo] == pkey_id(key));
    }
    const uint64_t km = ballot(keep);
    const int kept = popc64(km);
    if (keep) W.cs[lanes_below(km)] = key;
    LM_WSYNC();
    if (kept) {
        for (uint32_t t = lane; t < cnt; t += 64) {
            uint64_t bk = pkey(W.ld0[t], W.li0[t]);
            int lo = 0, hi = kept;
            while (lo < hi) { int mid = (lo + hi) >> 1; if (W.cs[mid] < bk) lo = mid + 1; else hi = mid; }
            uint32_t pos = t + (uint32_t)lo;
            if (pos < K) { W.ld1[pos] = W.ld0[t]; W.li1[pos] = W.li0[t]; }
        }
        if (lane < kept) {
            const uint64_t ck = W.cs[lane];
            uint32_t lo = 0, hi = cnt;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (pkey(W.ld0[mid], W.li0[mid]) < ck) lo = mid + 1; else hi = mid;
            }
            uint32_t pos = (uint32_t)lane + lo;
            if (pos < K) { W.ld1[pos] = pkey_d(ck); W.li1[pos] = pkey_id(ck); }
        }
        LM_WSYNC();
        const uint32_t ncnt = min(K, cnt + (uint32_t)kept);
        for (uint32_t t = lane; t < ncnt; t += 64) { rkd[t] = W.ld1[t]; rid[t] = W.li1[t]; }
        if (lane == 0) P.cnt[r] = ncnt;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent");
    LM_WSYNC();
}

// Operand fragments: 8 k-values per lane per 16-wide step, as bf16 (16 B) or,
// in the FP8 trial, as OCP e4m3 bytes (8 B) of x / s with s a power of two
// (the MFMA sum is scaled back by s^2).
template <bool F8>
struct LmFrag;
template <>
struct LmFrag<false> {
    lm_bf16x8 v;
    __device__ void load(const void* xq, uint32_t row, uint32_t d, uint32_t k0, bool ok) {
        uint4 z = make_uint4(0, 0, 0, 0);
        if (ok) z = *(const uint4*)((const uint16_t*)xq + (size_t)row * d + k0);
        v = __builtin_bit_cast(lm_bf16x8, z);
    }
};
template <>
struct LmFrag<true> {
    long v;
    __device__ void load(const void* xq, uint32_t row, uint32_t d, uint32_t k0, bool ok) {
        uint64_t z = 0;
        if (ok) z = *(const uint64_t*)((const uint8_t*)xq + (size_t)row * d + k0);
        v = (long)z;
    }
};
__device__ __forceinline__ lm_f32x16 lm_mma(const LmFrag<false>& a, const LmFrag<false>& b, lm_f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a.v, b.v, c, 0, 0, 0);
}
__device__ __forceinline__ lm_f32x16 lm_mma(const LmFrag<true>& a, const LmFrag<true>& b, lm_f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_fp8_fp8(a.v, b.v, c, 0, 0, 0);
}

template <int MC, bool F8>
__global__ void __launch_bounds__(256) k_leaf_mfma(LeafParams P, const void* __restrict__ xq, float s2,
                                                   const float* __restrict__ marg) {
    extern __shared__ __align__(16) char lds[];
    const uint32_t leaf = blockIdx.x;
    if (leaf >= P.nleaves) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    LmWave W = lm_carve(lds + lm_wave_bytes(P.K) * w, P.K);
    const uint32_t m = P.leaf_size[leaf], K = P.K, d = P.d;
    const uint32_t* mem = P.perm + P.leaf_start[leaf];
    const uint32_t nrb = (m + 31) / 32;
    uint32_t nsv = 0;   // pairs this lane let through the filter
    for (uint32_t rb = (uint32_t)w; rb < nrb; rb += 4) {
        // A fragments of the block's 32 members
        const uint32_t am = rb * 32 + (lane & 31);
        const uint32_t arow = am < m ? mem[am] : U32_NONE;
        LmFrag<F8> afr[LM_KS];
#pragma unroll
        for (int s = 0; s < LM_KS; s++) {
            const uint32_t k0 = 16 * s + 8 * (lane >> 5);
            afr[s].load(xq, arow, d, k0, arow != U32_NONE && k0 < d);
        }
        uint32_t rid[16];
        float bnr[16], thrm[16];
        auto load_thr = [&]() {
#pragma unroll
            for (int v = 0; v < 16; v++) {
                float t = __builtin_huge_valf();
                if (rid[v] != U32_NONE && P.cnt[rid[v]] >= K) t = P.keys[(size_t)rid[v] * K + K - 1] + marg[rid[v]];
                thrm[v] = t;
            }
        };
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const uint32_t a = rb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            rid[v] = a < m ? mem[a] : U32_NONE;
            bnr[v] = rid[v] != U32_NONE ? P.nrm[rid[v]] : 0.0f;
        }
        load_thr();
        if (lane < 32) W.svc[lane] = 0;
        LM_WSYNC();
        auto flush = [&]() {
            for (int rl = 0; rl < 32; rl++) {
                if (W.svc[rl] == 0) continue;
                lm_flush_row<MC>(P, W, mem, rl, mem[rb * 32 + rl]);
            }
            load_thr();
        };
        for (uint32_t ct = 0; ct < nrb; ct++) {
            const uint32_t cm = ct * 32 + (lane & 31);
            const uint32_t crow = cm < m ? mem[cm] : U32_NONE;
            lm_f32x16 acc;
#pragma unroll
            for (int v = 0; v < 16; v++) acc[v] = 0.0f;
            LmFrag<F8> bfr[LM_KS];
#pragma unroll
            for (int s = 0; s < LM_KS; s++) {
                const uint32_t k0 = 16 * s + 8 * (lane >> 5);
                bfr[s].load(xq, crow, d, k0, crow != U32_NONE && k0 < d);
            }
#pragma unroll
            for (int s = 0; s < LM_KS; s++) acc = lm_mma(afr[s], bfr[s], acc);
            const float bc = crow != U32_NONE ? P.nrm[crow] : 0.0f;
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const uint32_t a = rb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const float approx = fmaf(-2.0f, acc[v] * s2, bnr[v] + bc);
                if (a < m && cm < m && cm != a && approx <= thrm[v]) {
                    const int rl = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                    uint32_t p = atomicAdd(&W.svc[rl], 1u);
                    W.sv[rl * LM_SV + p] = (uint16_t)cm;
                    nsv++;
                }
            }
            LM_WSYNC();
            // a row gains at most 32 survivors per tile: flush before one could overflow
            if (__any(lane < 32 && W.svc[lane] > LM_SV - 32)) flush();
        }
        flush();
    }
    if (P.nsurv) {
        for (int o = 32; o > 0; o >>= 1) nsv += __shfl_xor(nsv, o);
        if (lane == 0 && nsv) atomicAdd(P.nsurv, (unsigned long long)nsv);
    }
}

size_t leaf_mfma_lds(uint32_t K) { return 4 * lm_wave_bytes(K); }

bool leaf_mfma_supported(uint32_t d, uint32_t max_leaf) {
    const uint32_t mc = d >> 4;
    return d <= 16 * LM_KS && d % 8 == 0 && mc >= 6 && mc <= 8 && max_leaf <= 65535;
}

template <int MC, bool F8>
static hipError_t launch_leaf_mfma_t(const LeafParams& P, const void* xq, float s2, const float* marg,
                                     hipStream_t s) {
    size_t lds = leaf_mfma_lds(P.K);
    hipFuncSetAttribute((const void*)k_leaf_mfma<MC, F8>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_leaf_mfma<MC, F8>), dim3(P.nleaves), dim3(256), lds, s, P, xq, s2, marg);
    return hipGetLastError();
}

template <bool F8>
static hipError_t launch_leaf_mfma_f(const LeafParams& P, const void* xq, float s2, const float* marg,
                                     hipStream_t s) {
    switch (P.d >> 4) {
        case 8: return launch_leaf_mfma_t<8, F8>(P, xq, s2, marg, s);
        case 7: return launch_leaf_mfma_t<7, F8>(P, xq, s2, marg, s);
        default: return launch_leaf_mfma_t<6, F8>(P, xq, s2, marg, s);
    }
}

hipError_t launch_leaf_mfma(const LeafParams& P, const uint16_t* xb, const float* marg, hipStream_t s) {
    if (P.nleaves == 0) return hipSuccess;
    if (!leaf_mfma_supported(P.d, P.max_leaf)) return hipErrorInvalidValue;
    return launch_leaf_mfma_f<false>(P, xb, 1.0f, marg, s);
}

hipError_t launch_leaf_mfma_fp8(const LeafParams& P, const uint8_t* x8, float scale, const float* marg,
                                hipStream_t s) {
    if (P.nleaves == 0) return hipSuccess;
    if (!leaf_mfma_supported(P.d, P.max_leaf)) return hipErrorInvalidValue;
    return launch_leaf_mfma_f<true>(P, x8, scale * scale, marg, s);
}

// x8 = e4m3(x / scale) (OCP FP8, round to nearest even); err[i] = |x_i - scale * x8_i|
// (rounded up), nx[i] = sqrt(nrm[i]) (rounded up): the inputs of the per-row
// filter bound for the FP8 sweep.
__global__ void k_to_fp8(const float* __restrict__ x, const float* __restrict__ nrm, uint32_t n, uint32_t d,
                         float inv_scale, float scale, uint8_t* __restrict__ x8, float* __restrict__ nx,
                         float* __restrict__ ex) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = x + (size_t)i * d;
    uint8_t* o = x8 + (size_t)i * d;
    float e2 = 0.0f;
    for (uint32_t k = 0; k < d; k += 2) {
        const float a = r[k] * inv_scale, b = k + 1 < d ? r[k + 1] * inv_scale : 0.0f;
        const int pk = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
        o[k] = (uint8_t)(pk & 0xff);
        if (k + 1 < d) o[k + 1] = (uint8_t)((pk >> 8) & 0xff);
        const float qa = __builtin_amdgcn_cvt_f32_fp8(pk, 0) * scale;
        const float da = r[k] - qa;
        e2 = fmaf(da, da, e2);
        if (k + 1 < d) {
            const float qb = __builtin_amdgcn_cvt_f32_fp8(pk, 1) * scale;
            const float db = r[k + 1] - qb;
            e2 = fmaf(db, db, e2);
        }
    }
    nx[i] = sqrtf(fmaxf(nrm[i], 0.0f)) * 1.0001f + 1e-30f;
    ex[i] = sqrtf(e2) * 1.001f + 1e-30f;
}

hipError_t launch_to_fp8(const float* x, const float* nrm, uint32_t n, uint32_t d, float scale, uint8_t* x8,
                         float* nx, float* ex, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_to_fp8, dim3((n + 255) / 256), dim3(256), 0, s, x, nrm, n, d, 1.0f / scale, scale, x8, nx,
                       ex);
    return hipGetLastError();
}
