// k_exact_mfma.hip -- exact self-excluded top-K (tools/dataset_gt.cpp:206-245)
// with the pair sweep on the matrix cores.
//
// Output is identical to the FP32 kernel (k_exact.hip): the K smallest
// (key, id) pairs with key = bn_i + bn_j - 2 <x_i, x_j> under the reference's
// FP32 contract.  The MFMA pass only decides which pairs need that exact key:
//
//   approx_ij = fmaf(-2, ip_bf16_ij, bn_i + bn_j)     (v_mfma_f32_32x32x16_bf16)
//   a pair can enter row i's top-K only if exact_ij <= tau_i (the row's
//   current K-th key), and |exact_ij - approx_ij| <= M_i, so pairs with
//   approx_ij > tau_i + M_i are skipped; the rest are re-scored exactly.
//
// M_i bounds the bf16 rounding of both operands, the FP32 accumulation error
// of the MFMA sum and of the reference's own FP32 dot product, and the final
// roundings (computed by k_exact_margin; see there).
//
// Tiling: a block owns 64 rows (4 wavefronts).  Column tiles of 128 rows of a
// bf16 copy of the base are staged in LDS (the next tile is prefetched into
// registers while the current one is swept).  Wavefront w computes rows
// 32 (w >> 1) .. +32 against tile columns 64 (w & 1) .. +64 as two 32x32 MFMA
// accumulators; the operand fragments follow the instruction's lane layout
// (A: row l % 32, k 8 (l / 32) .. +8; B: column l % 32, same k; C element v:
// row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32).  Surviving column ids
// collect in a per-row LDS buffer across tiles; only when some row's buffer
// could not take another tile (or after the last tile) do the wavefronts
// re-score their rows' buffers with the FP32 contract and merge those beating
// the row's K-th pair into its sorted list.  Between merges the filter keeps
// the thresholds of the last merge, which are never below the current ones,
// so it only lets more pairs through; batching the merges this way takes the
// per-survivor merge latency off the sweep.
#include "common.hpp"
#include "kernels.hpp"

typedef __bf16 xm_bf16x8 __attribute__((ext_vector_type(8)));
typedef float xm_f32x16 __attribute__((ext_vector_type(16)));

#define XM_C 128                     // columns per tile
#define XM_KS 8                      // 16-wide k steps: d <= 128, zero padded
#define XM_BS (16 * XM_KS + 8)       // bf16 per staged tile row (272 bytes)
#define XM_TV (XM_C * 16 * XM_KS / 8)   // 16-byte vectors per tile (2048)

// A block is RG groups of 32 rows (2 RG wavefronts, wavefront w: rows
// 32 (w >> 1) .., tile columns 64 (w & 1) ..): RG = 2 (64 rows) or 4 (128
// rows, when the LDS holds their lists).
struct XmLds {
    uint16_t* btile;   // XM_C x XM_BS bf16
    float* bnc;        // XM_C column norms
    float* ld;         // R x K (or, with the lists in global memory, waves x K)
    uint32_t* li;      // R x K (or waves x K)
    uint32_t* surv;    // R x sc column ids (survivors since the last merge)
    uint32_t* scnt;    // R
    uint32_t* flag;    // set when some row's buffer could not take another tile
    uint32_t* lcnt;    // R
    uint64_t* thr;     // R: K-th (key, id) pair, ~0 while the list is short
    float* thrm;       // R: K-th key + M_i (filter bound), +inf while short
    float* sd;         // waves x K merge scratch
    uint32_t* si;      // waves x K
    uint64_t* cs;      // waves x 64 sorted new pairs
    uint32_t* xid;     // waves x 64 column ids under exact evaluation
    float* xd;         // waves x 64 exact keys
};

__host__ __device__ inline size_t xm_r16(size_t b) { return (b + 15) & ~size_t(15); }

// gl: the rows' lists live in the output arrays (global memory) and a merge
// stages its row's list in a per-wavefront LDS copy (ld / li: waves x K).
__host__ __device__ inline size_t xm_carve(char* base, uint32_t R, uint32_t K, uint32_t sc, bool gl, XmLds* s) {
    const uint32_t NW = R / 16;
    size_t o = 0;
    auto take = [&](size_t bytes) { char* p = base ? base + o : nullptr; o += xm_r16(bytes); return p; };
    char* p;
    p = take(2ull * XM_C * XM_BS); if (s) s->btile = (uint16_t*)p;
    p = take(4ull * XM_C); if (s) s->bnc = (float*)p;
    p = take(4ull * (gl ? NW : R) * K); if (s) s->ld = (float*)p;
    p = take(4ull * (gl ? NW : R) * K); if (s) s->li = (uint32_t*)p;
    p = take(4ull * R * sc); if (s) s->surv = (uint32_t*)p;
    p = take(4ull * R); if (s) s->scnt = (uint32_t*)p;
    p = take(16); if (s) s->flag = (uint32_t*)p;
    p = take(4ull * R); if (s) s->lcnt = (uint32_t*)p;
    p = take(8ull * R); if (s) s->thr = (uint64_t*)p;
    p = take(4ull * R); if (s) s->thrm = (float*)p;
    p = take(4ull * NW * K); if (s) s->sd = (float*)p;
    p = take(4ull * NW * K); if (s) s->si = (uint32_t*)p;
    p = take(8ull * NW * 64); if (s) s->cs = (uint64_t*)p;
    p = take(4ull * NW * 64); if (s) s->xid = (uint32_t*)p;
    p = take(4ull * NW * 64); if (s) s->xd = (float*)p;
    return o;
}

#define XM_WSYNC()                                              \
    do {                                                        \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  \
        __builtin_amdgcn_wave_barrier();                        \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");  \
    } while (0)

// Merges the wave's sorted new pairs (key per lane, bc of them, ascending,
// all below the row's K-th pair) into row r's sorted list: in LDS, or (gk /
// gi non-null) in global memory, staged through the wavefront's LDS copy.
__device__ __forceinline__ void xm_merge(XmLds& S, int w, int r, uint32_t K, uint64_t key, int bc,
                                         float mrow, float* gk, uint32_t* gi) {
    const int lane = __lane_id();
    uint64_t* csw = S.cs + w * 64;
    csw[lane] = key;
    const uint32_t size = S.lcnt[r];
    float* rd = S.ld + (size_t)(gk ? w : r) * K;
    uint32_t* ri = S.li + (size_t)(gk ? w : r) * K;
    if (gk)
        for (uint32_t i = lane; i < size; i += 64) { rd[i] = gk[i]; ri[i] = gi[i]; }
    XM_WSYNC();
    float* od = S.sd + (size_t)w * K;
    uint32_t* oi = S.si + (size_t)w * K;
    for (uint32_t i = lane; i < size; i += 64) {
        uint64_t bk = pkey(rd[i], ri[i]);
        int lo = 0, hi = bc;
        while (lo < hi) { int mid = (lo + hi) >> 1; if (csw[mid] < bk) lo = mid + 1; else hi = mid; }
        uint32_t pos = i + (uint32_t)lo;
        if (pos < K) { od[pos] = rd[i]; oi[pos] = ri[i]; }
    }
    if (lane < bc) {
        uint32_t lo = 0, hi = size;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (pkey(rd[mid], ri[mid]) < key) lo = mid + 1; else hi = mid;
        }
        uint32_t pos = (uint32_t)lane + lo;
        if (pos < K) { od[pos] = pkey_d(key); oi[pos] = pkey_id(key); }
    }
    XM_WSYNC();
    const uint32_t ns = min(K, size + (uint32_t)bc);
    for (uint32_t i = lane; i < ns; i += 64) { rd[i] = od[i]; ri[i] = oi[i]; }
    if (gk) {
        for (uint32_t i = lane; i < ns; i += 64) { gk[i] = od[i]; gi[i] = oi[i]; }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");   // the next merge of the row reads them back
    }
    XM_WSYNC();
    if (lane == 0) {
        S.lcnt[r] = ns;
        if (ns >= K) {
            S.thr[r] = pkey(rd[K - 1], ri[K - 1]);
            S.thrm[r] = rd[K - 1] + mrow;
        }
    }
    XM_WSYNC();
}

// IPM: the inner-product key -<x_i, x_j> (dataset_gt.cpp:225); approx = -ip_bf16,
// and M_i (twice a bound on the inner product's error) still covers it.
// GL: the rows' lists in the output arrays (large K at 128-row blocks).
template <int MC, bool IPM, int RG, bool GL>
__global__ void __launch_bounds__(128 * RG) k_exact_mfma(const float* __restrict__ x,
                                                         const uint16_t* __restrict__ xb,
                                                         const float* __restrict__ nrm,
                                                         const float* __restrict__ marg, uint32_t n, uint32_t d,
                                                         uint32_t K, uint32_t sc, uint32_t row0, uint32_t rows,
                                                         uint32_t* __restrict__ out_ids,
                                                         float* __restrict__ out_keys) {
    constexpr int R = 32 * RG, NT = 128 * RG, NV = XM_TV / NT;
    extern __shared__ __align__(16) char lds[];
    XmLds S;
    xm_carve(lds, R, K, sc, GL, &S);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, l = lane & 15, g = lane >> 4;
    const uint32_t rb = row0 + blockIdx.x * R;
    const uint32_t rend = min(row0 + rows, rb + R);
    for (int r = tid; r < R; r += NT) {
        S.scnt[r] = 0;
        S.lcnt[r] = 0;
        S.thr[r] = ~0ull;
        S.thrm[r] = __builtin_huge_valf();
    }
    if (tid == 0) *S.flag = 0u;
    // A fragments: this wave's 32 rows, bf16, zero padded past d
    const int h = w >> 1, ch = w & 1;
    const uint32_t arow = rb + 32 * h + (lane & 31);
    xm_bf16x8 afr[XM_KS];
#pragma unroll
    for (int s = 0; s < XM_KS; s++) {
#pragma unroll
        for (int e = 0; e < 8; e++) {
            uint32_t k = 16 * s + 8 * (lane >> 5) + e;
            uint16_t b = (arow < rend && k < d) ? xb[(size_t)arow * d + k] : (uint16_t)0;
            afr[s][e] = __builtin_bit_cast(__bf16, b);
        }
    }
    // per-lane accumulator rows: v -> row 8 (v / 4) + 4 (l / 32) + v % 4 of the group
    float bnr[16];
    uint32_t rid[16];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        uint32_t rl = 32 * h + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        rid[v] = rb + rl;
        bnr[v] = rid[v] < rend ? nrm[rid[v]] : 0.0f;
    }
    // prefetch of tile 0
    const uint32_t vrow = d / 8;              // 16-byte vectors per base row
    uint4 pf[NV];
    auto fetch = [&](uint32_t c0) {
#pragma unroll
        for (int u = 0; u < NV; u++) {
            uint32_t e = (uint32_t)u * NT + (uint32_t)tid;     // vector index in the tile
            uint32_t c = e / (16 * XM_KS / 8), kv = e % (16 * XM_KS / 8);
            uint4 z = make_uint4(0, 0, 0, 0);
            if (c0 + c < n && kv < vrow) z = *(const uint4*)(xb + (size_t)(c0 + c) * d + 8 * kv);
            pf[u] = z;
        }
    };
    fetch(0);
    __syncthreads();
    for (uint32_t c0 = 0; c0 < n; c0 += XM_C) {
        // stage the prefetched tile, then prefetch the next one
#pragma unroll
        for (int u = 0; u < NV; u++) {
            uint32_t e = (uint32_t)u * NT + (uint32_t)tid;
            uint32_t c = e / (16 * XM_KS / 8), kv = e % (16 * XM_KS / 8);
            *(uint4*)(S.btile + (size_t)c * XM_BS + 8 * kv) = pf[u];
        }
        for (uint32_t c = tid; c < XM_C; c += NT) S.bnc[c] = (c0 + c < n) ? nrm[c0 + c] : 0.0f;
        __syncthreads();
        if (c0 + XM_C < n) fetch(c0 + XM_C);
        // sweep: two 32x32 accumulators per wave, their MFMA chains interleaved
        float thrm[16];
#pragma unroll
        for (int v = 0; v < 16; v++) thrm[v] = S.thrm[32 * h + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3)];
        xm_f32x16 acc[2];
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int v = 0; v < 16; v++) acc[t][v] = 0.0f;
        const uint16_t* bcol0 = S.btile + (size_t)(64 * ch + (lane & 31)) * XM_BS + 8 * (lane >> 5);
        const uint16_t* bcol1 = bcol0 + 32 * XM_BS;
#pragma unroll
        for (int s = 0; s < XM_KS; s++) {
            const xm_bf16x8 b0 = *(const xm_bf16x8*)(bcol0 + 16 * s);
            const xm_bf16x8 b1 = *(const xm_bf16x8*)(bcol1 + 16 * s);
            acc[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afr[s], b0, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(afr[s], b1, acc[1], 0, 0, 0);
        }
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int cl = 64 * ch + 32 * t + (lane & 31);     // tile column of this lane
            const uint32_t cid = c0 + (uint32_t)cl;
            const float bc = S.bnc[cl];
#pragma unroll
            for (int v = 0; v < 16; v++) {
                float approx = IPM ? -acc[t][v] : fmaf(-2.0f, acc[t][v], bnr[v] + bc);
                if (cid < n && rid[v] < rend && cid != rid[v] && approx <= thrm[v]) {
                    int rl = 32 * h + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
                    // below sc - XM_C before the tile, at most XM_C more: p < sc
                    uint32_t p = atomicAdd(&S.scnt[rl], 1u);
                    S.surv[(size_t)rl * sc + p] = cid;
                    if (p + 1u > sc - XM_C) *S.flag = 1u;
                }
            }
        }
        __syncthreads();
        if (!*S.flag && c0 + XM_C < n) continue;
        // exact re-scoring and merge: wave w owns rows 16 w .. 16 w + 15
        for (int rr = 0; rr < 16; rr++) {
            const int r = w * 16 + rr;
            const uint32_t cnt = S.scnt[r];
            if (cnt == 0) continue;
            const uint32_t row = rb + (uint32_t)r;
            QRow QR;
            qrow_load(QR, x + (size_t)row * d, (int)d, l);
            const float nr = nrm[row];
            const float mrow = marg[row];
            uint32_t* xid = S.xid + w * 64;
            float* xd = S.xd + w * 64;
            for (uint32_t b0 = 0; b0 < cnt; b0 += 64) {
                const int k = (int)min(64u, cnt - b0);
                if (lane < k) xid[lane] = S.surv[(size_t)r * sc + b0 + lane];
                XM_WSYNC();
                rows_batch<MC, IPM ? 3 : 1>(QR, x, d, nrm, nr, xid, k, l, g, xd);
                XM_WSYNC();
                const uint64_t th = S.thr[r];
                uint64_t key = ~0ull;
                if (lane < k) {
                    uint64_t pk = pkey(xd[lane], xid[lane]);
                    if (pk < th) key = pk;
                }
                key = wave_sort64(key);
                const int bc = popc64(ballot(key != ~0ull));
                if (bc) {
                    float* gk = GL ? out_keys + (size_t)(row - row0) * K : nullptr;
                    uint32_t* gi = GL ? out_ids + (size_t)(row - row0) * K : nullptr;
                    xm_merge(S, w, r, K, key, bc, mrow, gk, gi);
                }
            }
            if (lane == 0) S.scnt[r] = 0;
        }
        __syncthreads();
        if (tid == 0) *S.flag = 0u;   // ordered before the next filter by the next tile's barrier
    }
    for (int r = w * 16; r < w * 16 + 16; r++) {
        uint32_t row = rb + r;
        if (row >= rend) continue;
        for (uint32_t i = lane; i < K; i += 64) {
            if (GL) {
                if (i >= S.lcnt[r]) {
                    out_ids[(size_t)(row - row0) * K + i] = U32_NONE;
                    out_keys[(size_t)(row - row0) * K + i] = F32_BIG;
                }
            } else {
                out_ids[(size_t)(row - row0) * K + i] = i < S.lcnt[r] ? S.li[(size_t)r * K + i] : U32_NONE;
                out_keys[(size_t)(row - row0) * K + i] = i < S.lcnt[r] ? S.ld[(size_t)r * K + i] : F32_BIG;
            }
        }
    }
}

// Per-row filter margin M_i (see the header), evaluated in FP32 with every
// term rounded up by generous factors:
//   |approx - exact| <= 2 [ g ((|x|+e_x)(Y+E) + |x| Y) + e_x (Y+E) + |x| E ]
//                       + 4 u (|x| + Y)^2 + tiny
// with u = 2^-24, g = d u / (1 - d u) (any summation order), e_x = |x - bf16(x)|,
// Y = max |y|, E = max e_y over the columns.
__global__ void k_exact_norms2(const float* __restrict__ x, const float* __restrict__ nrm, uint32_t n, uint32_t d,
                               float* __restrict__ nx, float* __restrict__ ex) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = x + (size_t)i * d;
    float e2 = 0.0f;
    for (uint32_t k = 0; k < d; k++) {
        float v = r[k];
        uint32_t u = __float_as_uint(v);
        float b = __uint_as_float(((u + 0x7fffu + ((u >> 16) & 1u)) >> 16) << 16);
        float df = v - b;
        e2 = fmaf(df, df, e2);
    }
    nx[i] = sqrtf(fmaxf(nrm[i], 0.0f)) * 1.0001f + 1e-30f;
    ex[i] = sqrtf(e2) * 1.001f + 1e-30f;
}

__global__ void k_exact_margin(const float* __restrict__ nx, const float* __restrict__ ex, uint32_t n, uint32_t d,
                               float Y, float E, float* __restrict__ marg) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float u = 5.9604645e-08f;
    const float gm = (float)d * u / (1.0f - (float)d * u) * 1.01f;
    const float a = nx[i], e = ex[i];
    float m = 2.0f * (gm * ((a + e) * (Y + E) + a * Y) + e * (Y + E) + a * E) + 4.0f * u * (a + Y) * (a + Y);
    marg[i] = m * 1.01f + 1e-30f + 2.0f * (float)d * 1.2e-38f;
}

// Survivor buffer entries per row for R-row blocks (lists in LDS, or in
// global memory with gl): what the LDS left by the rest holds (a multiple of
// 64, at most 1024), or 0 when that is less than one tile plus 64.
static uint32_t exact_mfma_sc(uint32_t R, uint32_t K, bool gl) {
    const size_t budget = 160 * 1024 - 2048;
    const size_t rest = xm_carve(nullptr, R, K, 0, gl, nullptr);
    size_t sc = rest < budget ? (budget - rest) / (4 * (size_t)R) : 0;
    sc = std::min<size_t>(1024, sc) & ~size_t(63);
    return sc >= XM_C + 64 ? (uint32_t)sc : 0u;
}

// 128-row blocks (two wavefronts per SIMD): lists in LDS when they fit with
// a survivor buffer (small K), else in global memory (K <= 128 fits:
// 192 buffer entries per row); 64-row blocks with LDS lists otherwise.
struct XmShape {
    int rg;
    bool gl;
    uint32_t sc;
};
static XmShape exact_mfma_shape(uint32_t K) {
    if (uint32_t sc = exact_mfma_sc(128, K, false)) return {4, false, sc};
    if (uint32_t sc = exact_mfma_sc(128, K, true)) return {4, true, sc};
    return {2, false, exact_mfma_sc(64, K, false)};
}

size_t exact_mfma_lds(uint32_t K) {
    const XmShape h = exact_mfma_shape(K);
    return xm_carve(nullptr, 32u * (uint32_t)h.rg, K, h.sc, h.gl, nullptr);
}

bool exact_mfma_supported(uint32_t d, uint32_t K) {
    const uint32_t mc = d >> 4;
    return d <= 16 * XM_KS && d % 8 == 0 && mc >= 6 && mc <= 8 && K <= 128;
}

template <int MC, bool IPM, int RG, bool GL>
static hipError_t launch_exact_mfma_t(const float* x, const uint16_t* xb, const float* nrm, const float* marg,
                                      uint32_t n, uint32_t d, uint32_t K, uint32_t sc, uint32_t row0, uint32_t rows,
                                      uint32_t* ids, float* keys, hipStream_t s) {
    constexpr uint32_t R = 32 * RG;
    if (sc == 0) return hipErrorInvalidValue;
    size_t lds = xm_carve(nullptr, R, K, sc, GL, nullptr);
    if (hipError_t e = set_lds_limit((const void*)k_exact_mfma<MC, IPM, RG, GL>, lds); e != hipSuccess) return e;
    hipLaunchKernelGGL((k_exact_mfma<MC, IPM, RG, GL>), dim3((rows + R - 1) / R), dim3(128 * RG), lds, s, x, xb,
                       nrm, marg, n, d, K, sc, row0, rows, ids, keys);
    return hipGetLastError();
}

template <bool IPM, int RG, bool GL>
static hipError_t launch_exact_mfma_m(const float* x, const uint16_t* xb, const float* nrm, const float* marg,
                                      uint32_t n, uint32_t d, uint32_t K, uint32_t sc, uint32_t row0, uint32_t rows,
                                      uint32_t* ids, float* keys, hipStream_t s) {
    switch (d >> 4) {
        case 8: return launch_exact_mfma_t<8, IPM, RG, GL>(x, xb, nrm, marg, n, d, K, sc, row0, rows, ids, keys, s);
        case 7: return launch_exact_mfma_t<7, IPM, RG, GL>(x, xb, nrm, marg, n, d, K, sc, row0, rows, ids, keys, s);
        default: return launch_exact_mfma_t<6, IPM, RG, GL>(x, xb, nrm, marg, n, d, K, sc, row0, rows, ids, keys, s);
    }
}

template <bool IPM>
static hipError_t launch_exact_mfma_i(const float* x, const uint16_t* xb, const float* nrm, const float* marg,
                                      uint32_t n, uint32_t d, uint32_t K, uint32_t row0, uint32_t rows,
                                      uint32_t* ids, float* keys, hipStream_t s) {
    const XmShape h = exact_mfma_shape(K);
    if (h.rg == 4 && h.gl)
        return launch_exact_mfma_m<IPM, 4, true>(x, xb, nrm, marg, n, d, K, h.sc, row0, rows, ids, keys, s);
    if (h.rg == 4) return launch_exact_mfma_m<IPM, 4, false>(x, xb, nrm, marg, n, d, K, h.sc, row0, rows, ids, keys, s);
    return launch_exact_mfma_m<IPM, 2, false>(x, xb, nrm, marg, n, d, K, h.sc, row0, rows, ids, keys, s);
}

hipError_t launch_exact_mfma(const float* x, const uint16_t* xb, const float* nrm, const float* marg, uint32_t n,
                             uint32_t d, uint32_t K, uint32_t row0, uint32_t rows, int ipm, uint32_t* ids,
                             float* keys, hipStream_t s) {
    if (rows == 0) return hipSuccess;
    if (!exact_mfma_supported(d, K)) return hipErrorInvalidValue;
    return ipm ? launch_exact_mfma_i<true>(x, xb, nrm, marg, n, d, K, row0, rows, ids, keys, s)
               : launch_exact_mfma_i<false>(x, xb, nrm, marg, n, d, K, row0, rows, ids, keys, s);
}

hipError_t launch_exact_norms2(const float* x, const float* nrm, uint32_t n, uint32_t d, float* nx, float* ex,
                               hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_exact_norms2, dim3((n + 255) / 256), dim3(256), 0, s, x, nrm, n, d, nx, ex);
    return hipGetLastError();
}

hipError_t launch_exact_margin(const float* nx, const float* ex, uint32_t n, uint32_t d, float Y, float E,
                               float* marg, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_exact_margin, dim3((n + 255) / 256), dim3(256), 0, s, nx, ex, n, d, Y, E, marg);
    return hipGetLastError();
}
