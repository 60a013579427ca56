// k_leaf_mfma.hip -- RP-forest leaf all-pairs (init_knn, include/index.h:
// 1433-1449) with the pair sweep on the matrix cores, two passes per tree.
//
// Same result as k_leaf_knn (the FP32 sweep): after every tree, each row's
// sorted top-K list holds the K best (key, id) pairs among the pairs seen so
// far.  Exact keys use the reference's FP32 contract.
//
// Pass 1 (k_leaf_sweep): one block (4 wavefronts) per leaf.  A wavefront
// takes 32-member row blocks of the leaf and sweeps the leaf's 32-member
// column tiles with v_mfma_f32_32x32x16_bf16 (operands gathered from a bf16
// copy of the base through the tree permutation; the FP8 trial uses
// v_mfma_f32_32x32x16_fp8_fp8 on e4m3(x / s)).  A pair survives unless its
// approximate key exceeds the row's K-th key + M_r (M_r: the per-row bound of
// k_exact_margin); survivors are appended to the row's candidate list in HBM
// (the wave owns the row within this tree: every row sits in one leaf).
// The thresholds are those before the tree; a skipped pair has exact key >
// that K-th key >= every later K-th key, so it could never be listed.
//
// Pass 2 (k_leaf_merge): one wavefront per row with candidates: exact keys
// (16 lanes per candidate, four candidates per group in flight), pairs
// already listed dropped (identical pair -> identical key bits), the rest
// merged into the row's top-K held as a register list, the count reset.
#include "common.hpp"
#include "kernels.hpp"

typedef __bf16 lm_bf16x8 __attribute__((ext_vector_type(8)));
typedef float lm_f32x16 __attribute__((ext_vector_type(16)));

#define LM_KS 8     // 16-wide k steps: d <= 128, zero padded

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
__global__ void __launch_bounds__(256) k_leaf_sweep(LeafParams P, const void* __restrict__ xq, float s2,
                                                    const float* __restrict__ marg, uint32_t* __restrict__ cbuf,
                                                    uint32_t* __restrict__ ccnt, uint32_t capc) {
    __shared__ uint32_t gcnt[4][32];
    const uint32_t leaf = blockIdx.x;
    if (leaf >= P.nleaves) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t m = P.leaf_size[leaf], K = P.K, d = P.d;
    const uint32_t* mem = P.perm + P.leaf_start[leaf];
    const uint32_t nrb = (m + 31) / 32;
    uint32_t* gc = gcnt[w];
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
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const uint32_t a = rb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            rid[v] = a < m ? mem[a] : U32_NONE;
            bnr[v] = rid[v] != U32_NONE ? P.nrm[rid[v]] : 0.0f;
            float t = __builtin_huge_valf();
            if (rid[v] != U32_NONE && P.cnt[rid[v]] >= K) t = P.keys[(size_t)rid[v] * K + K - 1] + marg[rid[v]];
            thrm[v] = t;
        }
        if (lane < 32) gc[lane] = 0;
        wave_lds_sync();
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
                    const uint32_t slot = atomicAdd(&gc[rl], 1u);
                    cbuf[(size_t)rid[v] * capc + slot] = crow;
                    nsv++;
                }
            }
        }
        wave_lds_sync();
        if (lane < 32 && rb * 32 + lane < m) ccnt[mem[rb * 32 + lane]] = gc[lane];
        wave_lds_sync();
    }
    if (P.nsurv) {
        for (int o = 32; o > 0; o >>= 1) nsv += __shfl_xor(nsv, o);
        if (lane == 0 && nsv) atomicAdd(P.nsurv, (unsigned long long)nsv);
    }
}

// The same pass with the leaf staged in LDS (when it fits): the members'
// rows (bf16, or e4m3 bytes), ids, norms and filter bounds are copied once per
// leaf, so every A and B fragment of the sweep is an LDS read instead of a
// gather through the permutation.  Rows are padded to 16 LM_KS + 8 elements
// (the pad spreads the fragment reads over the banks); elements past d are
// zero.  Same survivors as k_leaf_sweep (the merge pass does not depend on
// their order).
#define LM_RS_EL (16 * LM_KS + 8)
template <bool F8>
__host__ __device__ constexpr uint32_t lm_row_bytes() { return LM_RS_EL * (F8 ? 1u : 2u); }
template <bool F8>
__host__ __device__ inline size_t lm_stage_lds(uint32_t max_leaf) {
    return (size_t)max_leaf * lm_row_bytes<F8>() + 12ull * max_leaf;
}

template <bool F8>
__device__ __forceinline__ void lm_frag_lds(LmFrag<F8>& f, const char* rows, uint32_t a, uint32_t k0, bool ok) {
    if constexpr (F8) {
        f.v = ok ? *(const long*)(rows + (size_t)a * lm_row_bytes<true>() + k0) : 0l;
    } else {
        uint4 z = make_uint4(0, 0, 0, 0);
        if (ok) z = *(const uint4*)(rows + (size_t)a * lm_row_bytes<false>() + 2 * k0);
        f.v = __builtin_bit_cast(lm_bf16x8, z);
    }
}

// 512 threads: two wavefronts per SIMD (one block fills a CU's LDS), so one
// wavefront's MFMA chain overlaps another's filter arithmetic.
#define LM_STAGE_THREADS 512
template <int MC, bool F8>
__global__ void __launch_bounds__(LM_STAGE_THREADS) k_leaf_sweep_lds(LeafParams P, const void* __restrict__ xq,
                                                                     float s2,
                                                        const float* __restrict__ marg, uint32_t* __restrict__ cbuf,
                                                        uint32_t* __restrict__ ccnt, uint32_t capc) {
    extern __shared__ __align__(16) char lds[];
    const uint32_t leaf = blockIdx.x;
    if (leaf >= P.nleaves) return;
    constexpr uint32_t ES = F8 ? 1u : 2u, RB = lm_row_bytes<F8>();
    constexpr uint32_t D = 16 * MC;                    // P.d
    const uint32_t m = P.leaf_size[leaf], K = P.K, ML = P.max_leaf;
    char* rows = lds;
    float* mnrm = (float*)(lds + (size_t)ML * RB);
    float* mthr = mnrm + ML;
    uint32_t* mid = (uint32_t*)(mthr + ML);

    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint32_t* mem = P.perm + P.leaf_start[leaf];
    // stage the members: rows (16-byte vectors), then the zero tail past d
    constexpr uint32_t VPR = D * ES / 16;
    // eight 16-byte loads in flight per thread before their LDS stores
    constexpr uint32_t U = 8;
    for (uint32_t e0 = 0; e0 < m * VPR; e0 += U * LM_STAGE_THREADS) {
        uint4 z[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t e = e0 + u * LM_STAGE_THREADS + (uint32_t)tid;
            z[u] = make_uint4(0, 0, 0, 0);
            if (e < m * VPR) z[u] = *(const uint4*)((const char*)xq + (size_t)mem[e / VPR] * D * ES + 16u * (e % VPR));
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t e = e0 + u * LM_STAGE_THREADS + (uint32_t)tid;
            if (e < m * VPR) *(uint4*)(rows + (size_t)(e / VPR) * RB + 16u * (e % VPR)) = z[u];
        }
    }
    if constexpr (D < 16 * LM_KS) {
        constexpr uint32_t TV = (16 * LM_KS - D) * ES / 16;     // 16-byte vectors of the zero tail
        for (uint32_t e = (uint32_t)tid; e < m * TV; e += LM_STAGE_THREADS) {
            const uint32_t a = e / TV, kv = e % TV;
            *(uint4*)(rows + (size_t)a * RB + D * ES + 16u * kv) = make_uint4(0, 0, 0, 0);
        }
    }
    for (uint32_t a = (uint32_t)tid; a < m; a += LM_STAGE_THREADS) {
        const uint32_t id = mem[a];
        mid[a] = id;
        mnrm[a] = P.nrm[id];
        float t = __builtin_huge_valf();
        if (P.cnt[id] >= K) t = P.keys[(size_t)id * K + K - 1] + marg[id];
        mthr[a] = t;
    }
    __syncthreads();
    const uint32_t nrb = (m + 31) / 32;
    uint32_t nsv = 0;   // pairs this lane let through the filter
    for (uint32_t rb = (uint32_t)w; rb < nrb; rb += LM_STAGE_THREADS / 64) {
        const uint32_t am = rb * 32 + (lane & 31);
        LmFrag<F8> afr[LM_KS];
#pragma unroll
        for (int s = 0; s < LM_KS; s++) lm_frag_lds<F8>(afr[s], rows, am, 16 * s + 8 * (lane >> 5), am < m);
        uint32_t rid[16];
        float bnr[16], thrm[16];
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const uint32_t a = rb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
            rid[v] = a < m ? mid[a] : U32_NONE;
            bnr[v] = a < m ? mnrm[a] : 0.0f;
            thrm[v] = a < m ? mthr[a] : -__builtin_huge_valf();
        }
        // survivor slots without atomics: accumulator element v of lane l
        // belongs to row 8 (v / 4) + 4 (l / 32) + v % 4, the same row for the
        // 32 lanes of a half-wavefront, so a ballot gives each passing lane
        // its slot and cnt[v] (uniform over the half) the row's count
        uint32_t cnt[16];
#pragma unroll
        for (int v = 0; v < 16; v++) cnt[v] = 0;
        const uint64_t half = (lane < 32) ? 0x00000000FFFFFFFFull : 0xFFFFFFFF00000000ull;
        const uint64_t below = (1ull << lane) - 1ull;
        auto filter = [&](uint32_t ct, const lm_f32x16& acc) __attribute__((always_inline)) {
            const uint32_t cm = ct * 32 + (lane & 31);
            const bool cok = cm < m;
            const float bc = cok ? mnrm[cm] : 0.0f;
            const uint32_t crow = cok ? mid[cm] : U32_NONE;
            const bool diag = ct == rb;   // the only tile holding self pairs
#pragma unroll
            for (int v = 0; v < 16; v++) {
                const uint32_t a = rb * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                const float approx = fmaf(-2.0f, F8 ? acc[v] * s2 : acc[v], bnr[v] + bc);
                const bool pass = cok && (!diag || cm != a) && approx <= thrm[v];   // thrm = -inf past m
                const uint64_t pm = ballot(pass) & half;
                if (pass) {
                    cbuf[(size_t)rid[v] * capc + cnt[v] + (uint32_t)popc64(pm & below)] = crow;
                    nsv++;
                }
                cnt[v] += (uint32_t)popc64(pm);
            }
        };
        // column tiles two at a time, their MFMA chains interleaved
        for (uint32_t ct = 0; ct < nrb; ct += 2) {
            const uint32_t cm0 = ct * 32 + (lane & 31), cm1 = cm0 + 32;
            lm_f32x16 acc0, acc1;
#pragma unroll
            for (int v = 0; v < 16; v++) { acc0[v] = 0.0f; acc1[v] = 0.0f; }
#pragma unroll
            for (int s = 0; s < LM_KS; s++) {
                LmFrag<F8> b0, b1;
                lm_frag_lds<F8>(b0, rows, cm0, 16 * s + 8 * (lane >> 5), cm0 < m);
                lm_frag_lds<F8>(b1, rows, cm1, 16 * s + 8 * (lane >> 5), cm1 < m);
                acc0 = lm_mma(afr[s], b0, acc0);
                acc1 = lm_mma(afr[s], b1, acc1);
            }
            filter(ct, acc0);
            if (ct + 1 < nrb) filter(ct + 1, acc1);
        }
        // row 8 (v / 4) + 4 h + v % 4 of the block: lane 32 h writes cnt[v]
        if ((lane & 31) == 0)
#pragma unroll
            for (int v = 0; v < 16; v++)
                if (rid[v] != U32_NONE) ccnt[rid[v]] = cnt[v];
    }
    if (P.nsurv) {
        for (int o = 32; o > 0; o >>= 1) nsv += __shfl_xor(nsv, o);
        if (lane == 0 && nsv) atomicAdd(P.nsurv, (unsigned long long)nsv);
    }
}

// Drops the lanes' candidate keys that equal a key of the list.
template <int NCH>
__device__ __forceinline__ uint64_t rk_drop_listed(const RKeys<NCH>& B, uint64_t key) {
    const int lane = __lane_id();
    uint64_t drop = 0;
    for (uint64_t mm = ballot(key != ~0ull); mm; mm &= mm - 1) {
        const int src = __ffsll((long long)mm) - 1;
        const uint64_t cv = readlane64(key, src);
        uint64_t hit = 0;
#pragma unroll
        for (int j = 0; j < NCH; j++) hit |= ballot(B.k[j] == cv);
        if (hit) drop |= 1ull << src;
    }
    return ((drop >> lane) & 1ull) ? ~0ull : key;
}

// Pass 2: rows r with ccnt[r] > 0; one wavefront (block of 64) per row.
template <int MC, int PCH>
__global__ void __launch_bounds__(64) k_leaf_merge(LeafParams P, uint32_t* __restrict__ cbuf,
                                                   uint32_t* __restrict__ ccnt, uint32_t capc) {
    __shared__ uint32_t tid[64];
    __shared__ float tdv[64];
    __shared__ uint64_t kb[64 * PCH];
    const uint32_t r = blockIdx.x;
    if (r >= P.n) return;
    const uint32_t c = ccnt[r];
    if (c == 0) return;
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    const uint32_t K = P.K, d = P.d;
    QRow QR;
    qrow_load(QR, P.x + (size_t)r * d, (int)d, l);
    const float nr = P.nrm[r];
    uint32_t* rid = P.ids + (size_t)r * K;
    float* rkd = P.keys + (size_t)r * K;
    uint32_t size = P.cnt[r];
    RKeys<PCH> B;
    B.x = 0;
#pragma unroll
    for (int j = 0; j < PCH; j++) {
        const uint32_t t = 64u * j + (uint32_t)lane;
        B.k[j] = t < size ? pkey(rkd[t], rid[t]) : ~0ull;
    }
    uint32_t added = 0;
    for (uint32_t b = 0; b < c; b += 64) {
        const int k = (int)min(64u, c - b);
        if (lane < k) tid[lane] = cbuf[(size_t)r * capc + b + lane];
        wave_lds_sync();
        rows_batch<MC, 1>(QR, P.x, d, P.nrm, nr, tid, k, l, g, tdv);
        wave_lds_sync();
        uint64_t key = lane < k ? pkey(tdv[lane], tid[lane]) : ~0ull;
        key = rk_drop_listed(B, key);
        added += rk_merge(B, key, size, K, kb);
    }
    if (added) {
#pragma unroll
        for (int j = 0; j < PCH; j++) {
            const uint32_t t = 64u * j + (uint32_t)lane;
            if (t < size) { rkd[t] = pkey_d(B.k[j]); rid[t] = pkey_id(B.k[j]); }
        }
        if (lane == 0) P.cnt[r] = size;
    }
    if (lane == 0) ccnt[r] = 0;
}

bool leaf_mfma_supported(uint32_t d, uint32_t max_leaf, uint32_t K) {
    const uint32_t mc = d >> 4;
    return d <= 16 * LM_KS && d % 8 == 0 && mc >= 6 && mc <= 8 && max_leaf <= 65535 && K <= 256;
}

template <int MC, bool F8>
static void leaf_sweep_t(const LeafParams& P, const void* xq, float s2, const float* marg, uint32_t* cbuf,
                         uint32_t* ccnt, uint32_t capc, hipStream_t s) {
    const size_t lds = lm_stage_lds<F8>(P.max_leaf);
    if (lds <= 160 * 1024 && set_lds_limit((const void*)k_leaf_sweep_lds<MC, F8>, lds) == hipSuccess) {
        hipLaunchKernelGGL((k_leaf_sweep_lds<MC, F8>), dim3(P.nleaves), dim3(LM_STAGE_THREADS), lds, s, P, xq, s2,
                           marg, cbuf, ccnt, capc);
        return;
    }
    hipLaunchKernelGGL((k_leaf_sweep<MC, F8>), dim3(P.nleaves), dim3(256), 0, s, P, xq, s2, marg, cbuf, ccnt, capc);
}

template <int MC, int PCH>
static void leaf_merge_t(const LeafParams& P, uint32_t* cbuf, uint32_t* ccnt, uint32_t capc, hipStream_t s) {
    hipLaunchKernelGGL((k_leaf_merge<MC, PCH>), dim3(P.n), dim3(64), 0, s, P, cbuf, ccnt, capc);
}

template <int MC>
static void leaf_pair_mc(const LeafParams& P, bool f8, const void* xq, float s2, const float* marg, uint32_t* cbuf,
                         uint32_t* ccnt, uint32_t capc, hipStream_t s) {
    if (f8) leaf_sweep_t<MC, true>(P, xq, s2, marg, cbuf, ccnt, capc, s);
    else leaf_sweep_t<MC, false>(P, xq, s2, marg, cbuf, ccnt, capc, s);
    if (P.K <= 128) leaf_merge_t<MC, 2>(P, cbuf, ccnt, capc, s);
    else leaf_merge_t<MC, 4>(P, cbuf, ccnt, capc, s);
}

hipError_t launch_leaf_mfma(const LeafParams& P, bool fp8, const void* xq, float scale, const float* marg,
                            uint32_t* cbuf, uint32_t* ccnt, uint32_t capc, hipStream_t s) {
    if (P.nleaves == 0) return hipSuccess;
    if (!leaf_mfma_supported(P.d, P.max_leaf, P.K) || capc < P.max_leaf) return hipErrorInvalidValue;
    const float s2 = fp8 ? scale * scale : 1.0f;
    switch (P.d >> 4) {
        case 8: leaf_pair_mc<8>(P, fp8, xq, s2, marg, cbuf, ccnt, capc, s); break;
        case 7: leaf_pair_mc<7>(P, fp8, xq, s2, marg, cbuf, ccnt, capc, s); break;
        default: leaf_pair_mc<6>(P, fp8, xq, s2, marg, cbuf, ccnt, capc, s); break;
    }
    return hipGetLastError();
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
