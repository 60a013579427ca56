// k_graph.hip -- lifted-graph and AKNN construction kernels.
//
//   warm start      include/index.h:1204-1218  (per-row sequential heap pushes)
//   LocalJoin       include/index.h:1468-1487  (snapshot schedule, see below)
//   reverse lists   include/heap.h:190-207     (count / scan / scatter)
//   gather          include/index.h:1491-1525  (one wavefront per row)
//   AKNN top-K      include/index.h:1348-1366  (pool sorted by (dist, id), self dropped)
//   prune           include/index.h:225-281, 772-829 (alpha-RNG to M)
//   commit          include/index.h:907-930
//
// Heap rows use the reference HeapList rule (include/heap.h:83-186): reject if
// value > root, reject duplicates, replace the root and sift down, so row
// storage order (which LocalJoin samples from) matches the reference rule.
//
// LocalJoin schedule.  The reference applies pushes under per-row locks in
// thread-arrival order, so its result depends on timing.  Here each pass
// takes every row's sample (first `sample` non-sentinel slots in storage
// order) at pass start, computes all pair distances, keeps proposals with
// d <= root of the target row (a pass-start snapshot; roots only decrease,
// so nothing that could be accepted is dropped), sorts them by
// (row, dist, id) and applies each row's proposals in that order with the
// heap rule.  The result is deterministic.
#include "common.hpp"
#include "kernels.hpp"
#include "edge_policy.hpp"

// ---- heap rule (one thread) ---------------------------------------------------
__device__ __forceinline__ int heap_push(uint32_t* idx, float* val, uint32_t cap, uint32_t node,
                                         float value) {
    if (value > val[0]) return 0;
    for (uint32_t s = 0; s < cap; s++)
        if (idx[s] == node) return 0;
    uint32_t cur = 0, sw;
    for (;;) {
        uint32_t l = 2 * cur + 1, r = l + 1;
        if (l >= cap) break;
        else if (r >= cap) { if (val[l] > value) sw = l; else break; }
        else if (val[l] >= val[r]) { if (val[l] > value) sw = l; else break; }
        else { if (val[r] > value) sw = r; else break; }
        idx[cur] = idx[sw];
        val[cur] = val[sw];
        cur = sw;
    }
    idx[cur] = node;
    val[cur] = value;
    return 1;
}

// d2(i, ids[i][t]) for every (row, slot); one 16-lane group per pair.
__global__ void __launch_bounds__(256) k_rows_d2(const float* __restrict__ x, const float* __restrict__ nrm,
                                                 uint32_t n, uint32_t d, const uint32_t* __restrict__ ids,
                                                 uint32_t width, float* __restrict__ out) {
    uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    int l = threadIdx.x & 15;
    if (t >= (uint64_t)n * width) return;
    uint32_t i = (uint32_t)(t / width), j = ids[t];
    if (j >= n || j == i) { if (l == 0) out[t] = F32_BIG; return; }
    float ip = g16_ip(x + (size_t)i * d, x + (size_t)j * d, (int)d, l);
    if (l == 0) out[t] = d2_combine(ip, nrm[i], nrm[j]);
}

__global__ void k_warm_push(const uint32_t* __restrict__ ids, const float* __restrict__ dv, uint32_t n,
                            uint32_t width, uint32_t cap, uint32_t* __restrict__ hidx,
                            float* __restrict__ hval) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t* hi = hidx + (size_t)i * cap;
    float* hv = hval + (size_t)i * cap;
    for (uint32_t s = 0; s < cap; s++) { hi[s] = U32_NONE; hv[s] = F32_BIG; }
    for (uint32_t t = 0; t < width; t++) {
        uint32_t j = ids[(size_t)i * width + t];
        if (j != i && j < n) heap_push(hi, hv, cap, j, dv[(size_t)i * width + t]);
    }
}

hipError_t launch_warm_start(const float* x, const float* nrm, uint32_t n, uint32_t d,
                             const uint32_t* ids, uint32_t width, uint32_t cap,
                             uint32_t* hidx, float* hval, hipStream_t s) {
    if (n == 0) return hipSuccess;
    float* dv = nullptr;
    hipError_t e = hipMallocAsync((void**)&dv, sizeof(float) * (size_t)n * width, s);
    if (e != hipSuccess) return e;
    uint64_t pairs = (uint64_t)n * width;
    hipLaunchKernelGGL(k_rows_d2, dim3((unsigned)((pairs * 16 + 255) / 256)), dim3(256), 0, s, x, nrm, n, d,
                       ids, width, dv);
    hipLaunchKernelGGL(k_warm_push, dim3((n + 127) / 128), dim3(128), 0, s, ids, dv, n, width, cap, hidx,
                       hval);
    e = hipGetLastError();
    const hipError_t ef = hipFreeAsync(dv, s);
    return e != hipSuccess ? e : ef;
}

// ---- LocalJoin ----------------------------------------------------------------------
__global__ void k_nnd_sample(const uint32_t* __restrict__ hidx, uint32_t n, uint32_t cap,
                             uint32_t sample, uint32_t* __restrict__ samp, uint32_t* __restrict__ scnt) {
    uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= n) return;
    uint32_t c = 0;
    for (uint32_t s = 0; s < cap && c < sample; s++) {
        uint32_t v = hidx[(size_t)u * cap + s];
        if (v != U32_NONE) samp[(size_t)u * sample + c++] = v;
    }
    scnt[u] = c;
}

// One wavefront per row u: sample vectors staged in LDS, all pairs (a < b).
__global__ void __launch_bounds__(64) k_nnd_propose(const float* __restrict__ x, const float* __restrict__ nrm,
                                                    uint32_t n, uint32_t d, const uint32_t* __restrict__ samp,
                                                    const uint32_t* __restrict__ scnt, uint32_t sample,
                                                    const float* __restrict__ hval, uint32_t cap,
                                                    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                    unsigned long long* __restrict__ count, uint64_t max_props) {
    extern __shared__ __align__(16) char lds[];
    float* sv = (float*)lds;                              // sample x d
    uint32_t* sid = (uint32_t*)(sv + (size_t)sample * d);  // sample
    float* snrm = (float*)(sid + sample);
    float* sroot = snrm + sample;
    uint32_t u = blockIdx.x;
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    uint32_t c = scnt[u];
    if (c < 2) return;
    for (uint32_t t = lane; t < c; t += 64) {
        uint32_t v = samp[(size_t)u * sample + t];
        sid[t] = v;
        snrm[t] = nrm[v];
        sroot[t] = hval[(size_t)v * cap];
    }
    __syncthreads();
    for (uint32_t e = lane; e < c * d; e += 64) {
        uint32_t t = e / d, k = e % d;
        sv[e] = x[(size_t)sid[t] * d + k];
    }
    __syncthreads();
    uint32_t npairs = c * (c - 1) / 2;
    for (uint32_t base = 0; base < npairs; base += 4) {
        uint32_t p = base + (uint32_t)g;
        int ok_vw = 0, ok_wv = 0;
        uint32_t a = 0, b = 0;
        float dd = 0.0f;
        if (p < npairs) {
            // p -> (a, b), a < b, row-major over the upper triangle
            uint32_t rem = p;
            a = 0;
            while (rem >= c - 1 - a) { rem -= c - 1 - a; a++; }
            b = a + 1 + rem;
            float ip = g16_ip(sv + (size_t)a * d, sv + (size_t)b * d, (int)d, l);
            dd = d2_combine(ip, snrm[a], snrm[b]);
            ok_vw = (l == 0) && !(dd > sroot[a]);   // push w into row v
            ok_wv = (l == 0) && !(dd > sroot[b]);   // push v into row w
        }
        uint64_t m1 = ballot(ok_vw), m2 = ballot(ok_wv);
        int tot = popc64(m1) + popc64(m2);
        if (tot == 0) continue;
        unsigned long long basep = 0;
        if (lane == 0) basep = atomicAdd(count, (unsigned long long)tot);
        basep = (unsigned long long)shfl64((uint64_t)basep, 0);
        if (ok_vw) {
            uint64_t pos = basep + lanes_below(m1);
            if (pos < max_props) { keys[pos] = ((uint64_t)sid[a] << 32) | fkey(dd); vals[pos] = sid[b]; }
        }
        if (ok_wv) {
            uint64_t pos = basep + popc64(m1) + lanes_below(m2);
            if (pos < max_props) { keys[pos] = ((uint64_t)sid[b] << 32) | fkey(dd); vals[pos] = sid[a]; }
        }
    }
}

// Chosen form of the LocalJoin proposals.  The sample's gram matrix (c <= 32
// rows, d <= 144) is one 32 x 32 bf16 MFMA tile per row u (the A and B
// fragments are the same rows); a pair is re-scored in FP32 (the plain
// kernel's bits) unless its approximate distance exceeds both endpoints'
// roots by more than the row bound marg (k_exact_margin over the lifted
// rows).  Proposals are staged in LDS and written with one counter update per
// row.  The proposal set equals the plain kernel's.
typedef __bf16 np_bf16x8 __attribute__((ext_vector_type(8)));
typedef float np_f32x16 __attribute__((ext_vector_type(16)));
#define NP_KS 9   // 16-wide k steps: d <= 144

__device__ __forceinline__ uint16_t np_bf16(float v) {   // round to nearest even (as k_exact_norms2)
    const uint32_t u = __float_as_uint(v);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

__global__ void __launch_bounds__(64) k_nnd_propose_mfma(const float* __restrict__ x, const float* __restrict__ nrm,
                                                         const float* __restrict__ marg, uint32_t n, uint32_t d,
                                                         const uint32_t* __restrict__ samp,
                                                         const uint32_t* __restrict__ scnt, uint32_t sample,
                                                         const float* __restrict__ hval, uint32_t cap,
                                                         uint64_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                         unsigned long long* __restrict__ count, uint64_t max_props) {
    extern __shared__ __align__(16) char lds[];
    const uint32_t npc = sample * (sample - 1) / 2;        // pair capacity
    uint64_t* pk = (uint64_t*)lds;                         // 2 npc proposals
    uint32_t* pv = (uint32_t*)(pk + 2 * npc);
    float* sv = (float*)(pv + 2 * npc);                    // sample x d
    uint32_t* sid = (uint32_t*)(sv + (size_t)sample * d);  // 32
    float* snrm = (float*)(sid + 32);
    float* sroot = snrm + 32;
    float* smg = sroot + 32;
    uint32_t* npair = (uint32_t*)(smg + 32);               // [0] pairs, [1] proposals
    uint32_t* pairs = npair + 2;                           // npc: a | b << 16
    const uint32_t u = blockIdx.x;
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    const uint32_t c = scnt[u];
    if (c < 2) return;
    if (lane < 32) {
        uint32_t v = (uint32_t)lane < c ? samp[(size_t)u * sample + lane] : 0u;
        sid[lane] = v;
        snrm[lane] = (uint32_t)lane < c ? nrm[v] : 0.0f;
        sroot[lane] = (uint32_t)lane < c ? hval[(size_t)v * cap] : 0.0f;
        smg[lane] = (uint32_t)lane < c ? marg[v] : 0.0f;
    }
    if (lane < 2) npair[lane] = 0;
    wave_lds_sync();
    for (uint32_t t = 0; t < c; t++) {
        const float* row = x + (size_t)sid[t] * d;
        for (uint32_t k = lane; k < d; k += 64) sv[(size_t)t * d + k] = row[k];
    }
    wave_lds_sync();
    // bf16 fragments of row l % 32 (A and B alike): k = 16 s + 8 (l / 32) .. +8
    const uint32_t ar = (uint32_t)(lane & 31);
    np_f32x16 acc;
#pragma unroll
    for (int v = 0; v < 16; v++) acc[v] = 0.0f;
#pragma unroll
    for (int s = 0; s < NP_KS; s++) {
        const uint32_t k0 = 16 * s + 8 * (lane >> 5);
        uint16_t h[8];
#pragma unroll
        for (int e = 0; e < 8; e++) h[e] = (ar < c && k0 + e < d) ? np_bf16(sv[(size_t)ar * d + k0 + e]) : (uint16_t)0;
        uint4 z;
        z.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
        z.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
        z.z = (uint32_t)h[4] | ((uint32_t)h[5] << 16);
        z.w = (uint32_t)h[6] | ((uint32_t)h[7] << 16);
        const np_bf16x8 f = __builtin_bit_cast(np_bf16x8, z);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(f, f, acc, 0, 0, 0);
    }
    // C element v of this lane: row a = 8 (v / 4) + 4 (l / 32) + v % 4, column b = l % 32
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const uint32_t a = 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3), b = ar;
        if (a < b && b < c) {
            const float approx = fmaf(-2.0f, acc[v], snrm[a] + snrm[b]);
            if (approx <= fmaxf(sroot[a], sroot[b]) + smg[a]) {
                const uint32_t slot = atomicAdd(&npair[0], 1u);
                pairs[slot] = a | (b << 16);
            }
        }
    }
    wave_lds_sync();
    const uint32_t np_ = npair[0];
    for (uint32_t base = 0; base < np_; base += 4) {
        const uint32_t p = base + (uint32_t)g;
        uint32_t a = 0, b = 0;
        float dd = 0.0f;
        if (p < np_) {
            a = pairs[p] & 0xffffu;
            b = pairs[p] >> 16;
            const float ip = g16_ip(sv + (size_t)a * d, sv + (size_t)b * d, (int)d, l);
            dd = d2_combine(ip, snrm[a], snrm[b]);
        }
        const int ok_vw = p < np_ && l == 0 && !(dd > sroot[a]);   // push w into row v
        const int ok_wv = p < np_ && l == 0 && !(dd > sroot[b]);   // push v into row w
        const uint64_t m1 = ballot(ok_vw), m2 = ballot(ok_wv);
        const uint32_t o = npair[1];
        if (ok_vw) {
            const uint32_t q = o + lanes_below(m1);
            pk[q] = ((uint64_t)sid[a] << 32) | fkey(dd);
            pv[q] = sid[b];
        }
        if (ok_wv) {
            const uint32_t q = o + popc64(m1) + lanes_below(m2);
            pk[q] = ((uint64_t)sid[b] << 32) | fkey(dd);
            pv[q] = sid[a];
        }
        wave_lds_sync();
        if (lane == 0) npair[1] = o + popc64(m1) + popc64(m2);
        wave_lds_sync();
    }
    const uint32_t tot = npair[1];
    if (tot == 0) return;
    unsigned long long basep = 0;
    if (lane == 0) basep = atomicAdd(count, (unsigned long long)tot);
    basep = (unsigned long long)shfl64((uint64_t)basep, 0);
    for (uint32_t q = lane; q < tot; q += 64) {
        const uint64_t pos = basep + q;
        if (pos < max_props) { keys[pos] = pk[q]; vals[pos] = pv[q]; }
    }
}

size_t nnd_propose_mfma_lds(uint32_t d, uint32_t sample) {
    const size_t npc = (size_t)sample * (sample - 1) / 2;
    return 12 * 2 * npc + 4ull * sample * d + 4 * 32 * 4 + 4 * 2 + 4 * npc;
}

bool nnd_propose_mfma_supported(uint32_t d, uint32_t sample) { return d <= 16 * NP_KS && sample >= 2 && sample <= 32; }

hipError_t launch_nnd_propose_mfma(const float* x, const float* nrm, const float* marg, uint32_t n, uint32_t d,
                                   const uint32_t* samp, const uint32_t* scnt, uint32_t sample, const float* hval,
                                   uint32_t cap, uint64_t* keys, uint32_t* vals, unsigned long long* count,
                                   uint64_t max_props, hipStream_t s) {
    if (n == 0) return hipSuccess;
    if (!nnd_propose_mfma_supported(d, sample)) return hipErrorInvalidValue;
    const size_t lds = nnd_propose_mfma_lds(d, sample);
    if (hipError_t e = set_lds_limit((const void*)k_nnd_propose_mfma, lds); e != hipSuccess) return e;
    hipLaunchKernelGGL(k_nnd_propose_mfma, dim3(n), dim3(64), lds, s, x, nrm, marg, n, d, samp, scnt, sample, hval,
                       cap, keys, vals, count, max_props);
    return hipGetLastError();
}

__global__ void k_row_bounds(const uint64_t* __restrict__ keys, uint64_t m, uint32_t n,
                             uint32_t* __restrict__ row_off) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > n) return;
    uint64_t target = (uint64_t)r << 32;
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
        uint64_t mid = (lo + hi) >> 1;
        if (keys[mid] < target) lo = mid + 1; else hi = mid;
    }
    row_off[r] = (uint32_t)lo;
}

// Applies each row's sorted proposals; equal distances go in ascending id order.
__global__ void k_nnd_apply(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                            const uint32_t* __restrict__ row_off, uint32_t n, uint32_t* __restrict__ hidx,
                            float* __restrict__ hval, uint32_t cap, unsigned long long* __restrict__ upd) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    uint32_t s = row_off[r], e = row_off[r + 1];
    if (s == e) return;
    uint32_t* hi = hidx + (size_t)r * cap;
    float* hv = hval + (size_t)r * cap;
    unsigned long long acc = 0;
    uint32_t t = s;
    while (t < e) {
        uint32_t run = t + 1;
        while (run < e && keys[run] == keys[t]) run++;
        float dd = fkey_inv((uint32_t)keys[t]);
        if (run == t + 1) {
            acc += heap_push(hi, hv, cap, vals[t], dd);
        } else {
            uint32_t last = 0;
            bool first = true;
            for (;;) {
                uint32_t best = U32_NONE;
                for (uint32_t q = t; q < run; q++) {
                    uint32_t v = vals[q];
                    if ((first || v > last) && v < best) best = v;
                }
                if (best == U32_NONE) break;
                acc += heap_push(hi, hv, cap, best, dd);
                last = best;
                first = false;
            }
        }
        t = run;
    }
    if (acc) atomicAdd(upd, acc);
}

hipError_t launch_nnd_sample(const uint32_t* hidx, uint32_t n, uint32_t cap, uint32_t sample,
                             uint32_t* samp, uint32_t* scnt, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nnd_sample, dim3((n + 255) / 256), dim3(256), 0, s, hidx, n, cap, sample, samp, scnt);
    return hipGetLastError();
}

hipError_t launch_nnd_propose(const float* x, const float* nrm, uint32_t n, uint32_t d,
                              const uint32_t* samp, const uint32_t* scnt, uint32_t sample,
                              const float* hval, uint32_t cap, uint64_t* keys, uint32_t* vals,
                              unsigned long long* count, uint64_t max_props, hipStream_t s) {
    if (n == 0) return hipSuccess;
    size_t lds = sizeof(float) * (size_t)sample * d + sizeof(uint32_t) * sample + 2 * sizeof(float) * sample;
    if (hipError_t e = set_lds_limit((const void*)k_nnd_propose, lds); e != hipSuccess) return e;
    hipLaunchKernelGGL(k_nnd_propose, dim3(n), dim3(64), lds, s, x, nrm, n, d, samp, scnt, sample, hval, cap,
                       keys, vals, count, max_props);
    return hipGetLastError();
}

hipError_t launch_row_bounds(const uint64_t* keys, uint64_t m, uint32_t n, uint32_t* row_off,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_row_bounds, dim3((n + 1 + 255) / 256), dim3(256), 0, s, keys, m, n, row_off);
    return hipGetLastError();
}

hipError_t launch_nnd_apply(const uint64_t* keys, const uint32_t* vals, uint64_t m,
                            const uint32_t* row_off, uint32_t n, uint32_t* hidx, float* hval,
                            uint32_t cap, unsigned long long* upd, hipStream_t s) {
    (void)m;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nnd_apply, dim3((n + 127) / 128), dim3(128), 0, s, keys, vals, row_off, n, hidx, hval,
                       cap, upd);
    return hipGetLastError();
}

// ---- reverse lists ---------------------------------------------------------------
__global__ void k_count_refs(const uint32_t* __restrict__ ids, uint64_t m, uint32_t n,
                             uint32_t* __restrict__ cnt) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    uint32_t j = ids[t];
    if (j < n) atomicAdd(&cnt[j], 1u);
}

// rows x width id table whose row r has rowcnt[r] valid leading entries
// (rowcnt == nullptr: sentinel-filtered full rows).
__global__ void k_scatter_rev(const uint32_t* __restrict__ ids, uint32_t rows, uint32_t width,
                              const uint32_t* __restrict__ rowcnt, uint32_t n,
                              uint32_t* __restrict__ cursor, uint32_t* __restrict__ rids) {
    uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)rows * width) return;
    uint32_t r = (uint32_t)(t / width), s = (uint32_t)(t % width);
    if (rowcnt && s >= rowcnt[r]) return;
    uint32_t j = ids[t];
    if (j >= n) return;
    uint32_t p = atomicAdd(&cursor[j], 1u);
    rids[p] = r;
}

hipError_t launch_count_refs(const uint32_t* ids, uint64_t m, uint32_t n, uint32_t* cnt, hipStream_t s) {
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_count_refs, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, ids, m, n, cnt);
    return hipGetLastError();
}

hipError_t launch_scatter_rev(const uint32_t* ids, uint32_t rows, uint32_t width,
                              const uint32_t* rowcnt, uint32_t n, uint32_t* cursor,
                              uint32_t* rids, hipStream_t s) {
    uint64_t m = (uint64_t)rows * width;
    if (m == 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_rev, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, ids, rows, width,
                       rowcnt, n, cursor, rids);
    return hipGetLastError();
}

// ---- sorted-list merge (one wavefront) ----------------------------------------------
// cs[0..kc) sorted candidate keys (LDS); (ad, ap) current list of size `size`
// (bit31 of ap = expanded).  Writes the first `cap` of the merge to (od, op).
__device__ __forceinline__ uint32_t merge_into(const uint64_t* cs, int kc, const float* ad,
                                               const uint32_t* ap, uint32_t size, uint32_t cap,
                                               float* od, uint32_t* op) {
    const int lane = __lane_id();
    for (uint32_t i = lane; i < size; i += 64) {
        uint64_t bk = pkey(ad[i], ap[i] & 0x7fffffffu);
        int lo = 0, hi = kc;
        while (lo < hi) { int mid = (lo + hi) >> 1; if (cs[mid] < bk) lo = mid + 1; else hi = mid; }
        uint32_t pos = i + (uint32_t)lo;
        if (pos < cap) { od[pos] = ad[i]; op[pos] = ap[i]; }
    }
    if (lane < kc) {
        uint64_t key = cs[lane];
        uint32_t lo = 0, hi = size;
        while (lo < hi) {
            uint32_t mid = (lo + hi) >> 1;
            if (pkey(ad[mid], ap[mid] & 0x7fffffffu) < key) lo = mid + 1; else hi = mid;
        }
        uint32_t pos = (uint32_t)lane + lo;
        if (pos < cap) { od[pos] = pkey_d(key); op[pos] = pkey_id(key); }
    }
    return min(cap, size + (uint32_t)kc);
}

// ---- gather (+ top-K or + prune) -------------------------------------------------------
struct GatherLds {
    float* qd0;       // beam queue, double buffered (selected by value so the
    float* qd1;       // struct stays in registers)
    uint32_t* qp0;
    uint32_t* qp1;
    __device__ float* qd(int b) const { return b ? qd1 : qd0; }
    __device__ uint32_t* qp(int b) const { return b ? qp1 : qp0; }
    uint32_t* hash;
    float* pd;        // pool (graph) or top-K (2 buffers in AKNN mode)
    uint32_t* pi;
    float* pd2;
    uint32_t* pi2;
    uint32_t* tid;
    float* tdv;
    uint64_t* cs;
    float* selv;      // prune: selected vectors (aliases the hash region)
    uint32_t* sel;
    char* pscratch;   // edge-policy prune scratch (after every other region)
};

__host__ __device__ __forceinline__ size_t rnd16(size_t b) { return (b + 15) & ~size_t(15); }

// Register-list form of the batched gather (queue <= 640 entries, AKNN
// top-K <= 128): LDS then holds only one exchange area per list.
__host__ __device__ __forceinline__ bool gather_rk(const GatherParams& P, bool graph) {
    return P.batched && P.explore <= 640 && (graph || P.K <= 128);
}

// Visited set: 32-bit entries (hmask + 1 slots); h16 1: 16-bit entries
// (ids < 2^20, H16_SLOTS_AKNN of them); h16 2: 16-bit entries plus the
// 32-bit overflow table (ids < 2^24).  A row whose first visits do not fit
// ends with status 1 and is redone on the global-memory path.
__host__ __device__ __forceinline__ size_t gather_hash_bytes(const GatherParams& P) {
    if (P.h16 == 2) return 2 * (size_t)H16_SLOTS + 4 * (size_t)H24_OVF_SLOTS;
    return P.h16 ? 2 * (size_t)H16_SLOTS_AKNN : 4 * (size_t)(P.hmask + 1);
}

// Bytes per entry of a register list's LDS exchange area.  With the 16-bit
// visited set the AKNN gather's lists move their keys as two 32-bit halves
// (rk_merge_x32), which halves the area and fits an eighth row per CU.
__host__ __device__ __forceinline__ size_t gather_xbytes(const GatherParams& P, bool graph) {
    return (!graph && P.h16) ? 4 : 8;
}

__host__ __device__ __forceinline__ size_t gather_region0(const GatherParams& P, bool graph, bool rk) {
    // hash + queue; in graph mode also the selected vectors of the prune step
    size_t a = rnd16(gather_hash_bytes(P)) + (rk ? rnd16(gather_xbytes(P, graph) * (size_t)P.explore) : 4 * rnd16(4 * (size_t)P.explore));
    size_t b = graph ? rnd16(4 * (size_t)P.M * P.d) + rnd16(4 * (size_t)P.M) : 0;
    return a > b ? a : b;
}

// Fresh-neighbour list of one expansion (batched form); a longer expansion is
// processed in several lists.  The AKNN gather's is shorter with the 16-bit
// visited set (its LDS area then fits more rows per CU).
#define GX_CAP_GRAPH 256
#define GX_CAP_AKNN 128
__host__ __device__ __forceinline__ uint32_t gather_tcap(const GatherParams& P, bool graph) {
    return P.batched ? ((graph || !P.h16) ? GX_CAP_GRAPH : GX_CAP_AKNN) : 64;
}

bool gather_register_lists(const GatherParams& P, bool graph) { return gather_rk(P, graph); }

size_t gather_lds_bytes(const GatherParams& P, bool graph, bool rk) {
    size_t pool = graph ? 2 * rnd16(4 * (size_t)P.pmax) : (rk ? rnd16(gather_xbytes(P, graph) * (size_t)P.K) : 4 * rnd16(4 * (size_t)P.K));
    return gather_region0(P, graph, rk) + pool + 2 * rnd16(4 * (size_t)gather_tcap(P, graph)) + (rk ? 0 : rnd16(8 * 64)) +
           ((graph && P.pol.on) ? rnd16(policy_scratch_bytes(std::max<uint32_t>(P.pmax, 64), P.pol, P.M)) : 0);
}

// rk: register lists; the queue's exchange area is qd0 (explore entries), the
// top-K's is pd (K entries), gather_xbytes each; the other list arrays are absent.
__device__ __forceinline__ GatherLds gather_carve(char* base, const GatherParams& P, bool graph, bool rk) {
    GatherLds s;
    size_t o = 0;
    auto take = [&](size_t bytes) { char* p = base + o; o += rnd16(bytes); return p; };
    s.hash = (uint32_t*)take(gather_hash_bytes(P));
    if (rk) {
        s.qd0 = (float*)take(gather_xbytes(P, graph) * (size_t)P.explore);
        s.qd1 = nullptr; s.qp0 = nullptr; s.qp1 = nullptr;
    } else {
        s.qd0 = (float*)take(4 * (size_t)P.explore);
        s.qd1 = (float*)take(4 * (size_t)P.explore);
        s.qp0 = (uint32_t*)take(4 * (size_t)P.explore);
        s.qp1 = (uint32_t*)take(4 * (size_t)P.explore);
    }
    s.selv = (float*)base;
    s.sel = (uint32_t*)(base + rnd16(4 * (size_t)P.M * P.d));
    o = gather_region0(P, graph, rk);
    if (graph) {
        s.pd = (float*)take(4 * (size_t)P.pmax);
        s.pi = (uint32_t*)take(4 * (size_t)P.pmax);
        s.pd2 = nullptr;
        s.pi2 = nullptr;
    } else if (rk) {
        s.pd = (float*)take(gather_xbytes(P, graph) * (size_t)P.K);
        s.pi = nullptr; s.pd2 = nullptr; s.pi2 = nullptr;
    } else {
        s.pd = (float*)take(4 * (size_t)P.K);
        s.pi = (uint32_t*)take(4 * (size_t)P.K);
        s.pd2 = (float*)take(4 * (size_t)P.K);
        s.pi2 = (uint32_t*)take(4 * (size_t)P.K);
    }
    s.tid = (uint32_t*)take(4 * (size_t)gather_tcap(P, graph));
    s.tdv = (float*)take(4 * (size_t)gather_tcap(P, graph));
    s.cs = rk ? nullptr : (uint64_t*)take(8 * 64);
    s.pscratch = (graph && P.pol.on) ? take(policy_scratch_bytes(max(P.pmax, 64u), P.pol, P.M)) : nullptr;
    return s;
}

// Bitonic sort of (pd, pi) [0, m) by (dist, id) in place; m <= cap2 (power of 2).
__device__ void lds_sort_pairs(float* pd, uint32_t* pi, uint32_t m, uint32_t cap2) {
    const int lane = __lane_id();
    for (uint32_t t = m + lane; t < cap2; t += 64) { pd[t] = F32_BIG; pi[t] = U32_NONE; }
    __syncthreads();
    for (uint32_t k = 2; k <= cap2; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = lane; t < cap2; t += 64) {
                uint32_t p = t ^ j;
                if (p > t) {
                    uint64_t a = pkey(pd[t], pi[t]), b = pkey(pd[p], pi[p]);
                    bool up = (t & k) == 0;
                    if ((a > b) == up) {
                        float fd = pd[t]; pd[t] = pd[p]; pd[p] = fd;
                        uint32_t fi = pi[t]; pi[t] = pi[p]; pi[p] = fi;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// alpha-RNG selection over a (dist, id)-sorted, duplicate-free pool
// (include/index.h:267-281); selected vectors are staged in LDS.  MC = d >> 4
// (or 0): each candidate row and each selected vector is requested whole
// before its dot product starts.
template <int MC>
__device__ uint32_t rng_select(const GatherParams& P, const float* pd, const uint32_t* pi, uint32_t m,
                               float* selv, uint32_t* sel, uint32_t* out) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    const int d = (int)P.d;
    uint32_t deg = 0;
    __shared__ float selnrm[64];
    for (uint32_t t = 0; t < m && deg < P.M; t++) {
        uint32_t cid = pi[t];
        float d2c = pd[t];
        QRow QC;
        qrow_load_mc<MC>(QC, P.x + (size_t)cid * d, d, l);
        float nc = P.nrm[cid];
        int occ = 0;
        for (uint32_t s0 = 0; s0 < deg; s0 += 4) {
            uint32_t s = s0 + (uint32_t)g;
            int o = 0;
            if (s < deg) {
                float ip = qrow_ip_mc<MC>(QC, selv + (size_t)s * d, d, l);
                float dsc = d2_combine(ip, selnrm[s], nc);
                o = (P.alpha * dsc < d2c);
            }
            if (__any(o)) { occ = 1; break; }
        }
        if (!occ) {
            // stage the vector: group 0 writes its chunks
            if (g == 0) {
                int mm = d >> 4, r = d & 15;
#pragma unroll
                for (int c = 0; c < QMAX_CHUNKS; c++)
                    if (c < mm) selv[(size_t)deg * d + 16 * c + l] = QC.v[c];
                if (l < r) selv[(size_t)deg * d + 16 * mm + l] = QC.tail;
            }
            if (lane == 0) { sel[deg] = cid; selnrm[deg] = nc; out[deg] = cid; }
            deg++;
            __syncthreads();
        }
    }
    return deg;
}

// One row of the gather.  QCH > 0: the beam queue is a register list of
// QCH chunks (explore <= 64 QCH); PCH > 0: so is the AKNN top-K (K <= 64 PCH).
// Otherwise both live sorted in LDS (double buffered, binary-search merges).
// HM: the visited set (P.h16: 0 32-bit entries; 1, 2 16-bit entries, with the
// AKNN lists' 32-bit exchange).
template <bool GRAPH, bool GWS, int MC, int QCH, int PCH, int HM>
__device__ __forceinline__ void gather_row(const GatherParams& P, GatherLds S, uint32_t i) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    const int d = (int)P.d;
    if constexpr (HM != 0) {
        for (uint32_t j = lane; j < (HM == 1 ? H16_SLOTS_AKNN : H16_SLOTS) / 8u; j += 64)
            ((uint4*)S.hash)[j] = make_uint4(~0u, ~0u, ~0u, ~0u);
        if constexpr (HM == 2)
            for (uint32_t j = lane; j < H24_OVF_SLOTS; j += 64) S.hash[H16_SLOTS / 2u + j] = 0u;
    } else {
        for (uint32_t j = lane; j <= P.hmask; j += 64) S.hash[j] = 0u;
    }
    __syncthreads();
    QRow QI;
    qrow_load(QI, P.x + (size_t)i * d, d, l);
    const float ni = P.nrm[i];
    uint32_t qsize = 0, psize = 0, pcount = 0;
    uint64_t nd2 = 0, nadj = 0;   // work counts (distances computed, adjacency ids read)
    int qb = 0, pb = 0, status = 0;
    RKeys<(QCH > 0 ? QCH : 1)> Q;
    RKeys<(PCH > 0 ? PCH : 1)> PL;
    rk_clear(Q);
    rk_clear(PL);

    // visit: a chunk of candidate ids (lane-held, in reference order) goes
    // through the visited set; first visits are appended to tid[nf..] (with
    // seed_d, the stored distance, for the seeds).
    // process: distances for tid[0..nf) (unless seeded), then pool and queue
    // merges, 64 candidates at a time.  Merging a set of candidates in pieces
    // or at once gives the same sorted, truncated pool and queue (total
    // (dist, id) order, no duplicates), so one expansion's first visits can
    // be processed together: the stop test (pool count) is only taken between
    // expansions.
    uint32_t nf = 0;
    auto visit = [&](uint32_t id, int valid, float seed_d) __attribute__((always_inline)) -> int {
        int fresh = 0;
        if (valid) {
            int r;
            if constexpr (HM == 1) r = hset16_insert((uint16_t*)S.hash, H16_LB_AKNN, id);
            else if constexpr (HM == 2) r = hset24_insert((uint16_t*)S.hash, S.hash + H16_SLOTS / 2u, id);
            else r = hset_insert(S.hash, P.hmask, id);
            if (r < 0) status = 1;
            fresh = (r == 1);
        }
        if (__any(status)) { status = 1; return 0; }
        uint64_t fm = ballot(fresh);
        if (fresh) { S.tid[nf + lanes_below(fm)] = id; S.tdv[nf + lanes_below(fm)] = seed_d; }
        nf += (uint32_t)popc64(fm);
        return 1;
    };
    auto process = [&](bool seeded) __attribute__((always_inline)) -> int {
        const uint32_t kall = nf;
        nf = 0;
        if (kall == 0) return 1;
        __syncthreads();
        if (!seeded) {
            nd2 += kall;
            if (!GRAPH && P.ipm) {
                // "ip" (AKNN source only): one first visit per lane
                for (uint32_t t = (uint32_t)lane; t < kall; t += 64)
                    S.tdv[t] = ipdist_seq(P.x + (size_t)i * d, P.x + (size_t)S.tid[t] * d, (uint32_t)d);
            } else if constexpr (MC > 0) {
                // register-list form: sixteen rows per group in flight for
                // the graph (about eight first visits per expansion), four
                // for the AKNN source (one or two), whose registers then
                // leave room for more rows per CU
                if constexpr (QCH > 0 && GRAPH) rows_batch<MC, 1, 16>(QI, P.x, (uint32_t)d, P.nrm, ni, S.tid, (int)kall, l, g, S.tdv);
                else if constexpr (QCH > 0) rows_batch<MC, 1, 4>(QI, P.x, (uint32_t)d, P.nrm, ni, S.tid, (int)kall, l, g, S.tdv);
                else rows_batch<MC, 1>(QI, P.x, (uint32_t)d, P.nrm, ni, S.tid, (int)kall, l, g, S.tdv);
            } else {
                for (uint32_t s = g; s < kall; s += 4) {
                    uint32_t c = S.tid[s];
                    float ip = qrow_ip(QI, P.x + (size_t)c * d, d, l);
                    if (l == 0) S.tdv[s] = d2_combine(ip, ni, P.nrm[c]);
                }
            }
            __syncthreads();
        }
        for (uint32_t b = 0; b < kall; b += 64) {
            const int k = (int)min(64u, kall - b);
            float dv = lane < k ? S.tdv[b + lane] : F32_BIG;
            uint32_t cid = lane < k ? S.tid[b + lane] : U32_NONE;
            // pool
            if (GRAPH) {
                if (psize + (uint32_t)k > P.pmax) { status = 2; return 0; }
                if (lane < k) { S.pd[psize + lane] = dv; S.pi[psize + lane] = cid; }
                psize += (uint32_t)k;
            } else if constexpr (PCH > 0) {
                if constexpr (HM != 0) rk_merge_x32(PL, (lane < k && cid != i) ? pkey(dv, cid) : ~0ull, psize, P.K, (uint32_t*)S.pd);
                else rk_merge(PL, (lane < k && cid != i) ? pkey(dv, cid) : ~0ull, psize, P.K, (uint64_t*)S.pd);
            } else {
                uint64_t key = (lane < k && cid != i) ? pkey(dv, cid) : ~0ull;
                float* ad = pb ? S.pd2 : S.pd;
                uint32_t* ap = pb ? S.pi2 : S.pi;
                if (psize == P.K && key != ~0ull && !(key < pkey(ad[P.K - 1], ap[P.K - 1]))) key = ~0ull;
                int kc = popc64(ballot(key != ~0ull));
                if (kc) {
                    key = wave_sort64(key);
                    S.cs[lane] = key;
                    __syncthreads();
                    psize = merge_into(S.cs, kc, ad, ap, psize, P.K, pb ? S.pd : S.pd2, pb ? S.pi : S.pi2);
                    pb ^= 1;
                    __syncthreads();
                }
            }
            pcount += (uint32_t)k;
            // queue
            if constexpr (QCH > 0) {
                if constexpr (GRAPH) rk_merge(Q, lane < k ? pkey(dv, cid) : ~0ull, qsize, P.explore, (uint64_t*)S.qd0);
                else if constexpr (HM != 0) rk_merge_x32(Q, lane < k ? pkey(dv, cid) : ~0ull, qsize, P.explore, (uint32_t*)S.qd0);
                else rk_merge(Q, lane < k ? pkey(dv, cid) : ~0ull, qsize, P.explore, (uint64_t*)S.qd0);
            } else {
                uint64_t key = lane < k ? pkey(dv, cid) : ~0ull;
                float* ad = S.qd(qb);
                uint32_t* ap = S.qp(qb);
                if (qsize == P.explore && key != ~0ull &&
                    !(key < pkey(ad[P.explore - 1], ap[P.explore - 1] & 0x7fffffffu)))
                    key = ~0ull;
                int kc = popc64(ballot(key != ~0ull));
                if (kc) {
                    key = wave_sort64(key);
                    S.cs[lane] = key;
                    __syncthreads();
                    qsize = merge_into(S.cs, kc, ad, ap, qsize, P.explore, S.qd(qb ^ 1), S.qp(qb ^ 1));
                    qb ^= 1;
                }
                __syncthreads();
            }
        }
        return 1;
    };
    const uint32_t tcap = gather_tcap(P, GRAPH);

    if (GRAPH && P.vmode) {
        // Vamana collect (index.h:831-880): the beam starts at the entry,
        // whose distance is computed like every later visit's
        const uint32_t ep = P.ventry ? P.ventry[i] : P.ventry0;
        if (visit(lane == 0 ? ep : U32_NONE, lane == 0 && ep < P.n, 0.0f)) process(false);
    } else {
        // seeds: the row's own slots in storage order, with stored values
        for (uint32_t s0 = 0; s0 < P.cap && !status; s0 += 64) {
            uint32_t s = s0 + lane;
            uint32_t id = s < P.cap ? P.hidx[(size_t)i * P.cap + s] : U32_NONE;
            float dv = s < P.cap ? P.hval[(size_t)i * P.cap + s] : F32_BIG;
            if (visit(id, id != U32_NONE && id < P.n, dv)) process(true);
        }
    }
    if (GRAPH && P.vmode) {
        // the beam expands rows of the fixed-degree graph (one neighbour per
        // lane, M <= 64) until no entry is left unexpanded
        while (!status) {
            uint32_t c;
            if constexpr (QCH > 0) {
                c = rk_take(Q, true);
                if (c == U32_NONE) break;
            } else {
                uint32_t cur = U32_NONE;
                for (uint32_t b = 0; b < qsize; b += 64) {
                    uint32_t t = b + lane;
                    uint64_t m = ballot(t < qsize && !(S.qp(qb)[t] >> 31));
                    if (m) { cur = b + (uint32_t)(__ffsll((long long)m) - 1); break; }
                }
                if (cur == U32_NONE) break;
                c = S.qp(qb)[cur];
                __syncthreads();
                if (lane == 0) S.qp(qb)[cur] = c | 0x80000000u;
                __syncthreads();
            }
            const uint32_t dg = P.vdeg[c];
            const uint32_t id = (uint32_t)lane < dg ? P.vlinks[(size_t)c * P.vM + lane] : U32_NONE;
            nadj += dg;
            if (visit(id, id != U32_NONE && id < P.n, 0.0f)) process(false);
        }
    } else if constexpr (QCH > 0) {
        // beam over the neighbour heaps and their reverse lists
        // Register-list form.  An expansion's adjacency (heap slots, the
        // reverse-list bounds and its first 64 entries) is requested at once,
        // and that of the entry next in line once c is marked (pc) is
        // fetched while c is expanded; it is used when no new candidate of
        // c overtakes it.  The visits keep the reference order (heap slots,
        // then the reverse list), and every first visit of an expansion is
        // processed before the stop test, as in the loop below.
        uint32_t pc = U32_NONE, pa0 = U32_NONE, pa1 = U32_NONE, pr0 = 0, pr1 = 0;
        const uint32_t cap = P.cap;
        auto heap_slot = [&](uint32_t node, uint32_t s) -> uint32_t {
            return s < cap ? P.hidx[(size_t)node * cap + s] : U32_NONE;
        };
        while (!status) {
            const uint32_t c = rk_take(Q, true);
            if (c == U32_NONE) break;
            uint32_t a0, a1, r0, r1;
            if (c == pc) {
                a0 = pa0; a1 = pa1; r0 = pr0; r1 = pr1;
            } else {
                a0 = heap_slot(c, (uint32_t)lane);
                a1 = heap_slot(c, 64u + (uint32_t)lane);
                r0 = P.roff[c];
                r1 = P.roff[c + 1];
            }
            const uint32_t rv0 = (r0 + (uint32_t)lane < r1) ? P.rids[r0 + (uint32_t)lane] : U32_NONE;
            pc = rk_take(Q, false);
            if (pc != U32_NONE) {
                pa0 = heap_slot(pc, (uint32_t)lane);
                pa1 = heap_slot(pc, 64u + (uint32_t)lane);
                pr0 = P.roff[pc];
                pr1 = P.roff[pc + 1];
            }
            // chunks t < nh: heap slots 64 t.. ; then the reverse list from r0
            const uint32_t nh = (cap + 63) / 64, nt = nh + (r1 - r0 + 63) / 64;
            nadj += cap + (r1 - r0);
            for (uint32_t t = 0; t < nt && !status; t++) {
                uint32_t id;
                if (t == 0) id = a0;
                else if (t == 1 && nh > 1) id = a1;
                else if (t < nh) id = heap_slot(c, 64u * t + (uint32_t)lane);
                else if (t == nh) id = rv0;
                else {
                    const uint32_t sr = r0 + 64u * (t - nh) + (uint32_t)lane;
                    id = sr < r1 ? P.rids[sr] : U32_NONE;
                }
                if (visit(id, id != U32_NONE && id < P.n, 0.0f) && nf + 64 > tcap) process(false);
            }
            if (!status) process(false);
            if (pcount >= P.pool_cap) break;
        }
    } else while (!status) {
        uint32_t c;
        {
            uint32_t cur = U32_NONE;
            for (uint32_t b = 0; b < qsize; b += 64) {
                uint32_t t = b + lane;
                uint64_t m = ballot(t < qsize && !(S.qp(qb)[t] >> 31));
                if (m) { cur = b + (uint32_t)(__ffsll((long long)m) - 1); break; }
            }
            if (cur == U32_NONE) break;
            c = S.qp(qb)[cur];
            __syncthreads();
            if (lane == 0) S.qp(qb)[cur] = c | 0x80000000u;
            __syncthreads();
        }
        for (uint32_t s0 = 0; s0 < P.cap && !status; s0 += 64) {
            uint32_t s = s0 + lane;
            uint32_t id = s < P.cap ? P.hidx[(size_t)c * P.cap + s] : U32_NONE;
            if (visit(id, id != U32_NONE && id < P.n, 0.0f) && (!P.batched || nf + 64 > tcap)) process(false);
        }
        uint32_t r0 = P.roff[c], r1 = P.roff[c + 1];
        nadj += P.cap + (r1 - r0);
        for (uint32_t s0 = r0; s0 < r1 && !status; s0 += 64) {
            uint32_t s = s0 + lane;
            uint32_t id = s < r1 ? P.rids[s] : U32_NONE;
            if (visit(id, id != U32_NONE && id < P.n, 0.0f) && (!P.batched || nf + 64 > tcap)) process(false);
        }
        // plain form: each chunk was processed on its own; batched form: the
        // expansion's first visits together (in lists of up to tcap)
        if (!status) process(false);
        if (pcount >= P.pool_cap) break;
    }
    if (GRAPH && P.vmode && !status) {
        // append_aknn_edges (index.h:766-769): the row's heap slots, whose
        // stored distances are the d2 the heap took them with; vmode 2 also
        // append_current_edges (:757-764).  Through the visited set: an id
        // already in the pool carries the same distance there.
        for (uint32_t s0 = 0; s0 < P.cap && !status; s0 += 64) {
            const uint32_t s = s0 + lane;
            const uint32_t id = s < P.cap ? P.hidx[(size_t)i * P.cap + s] : U32_NONE;
            const float dv = s < P.cap ? P.hval[(size_t)i * P.cap + s] : F32_BIG;
            if (visit(id, id != U32_NONE && id < P.n && id != i, dv)) process(true);
        }
        if (P.vmode == 2 && !status) {
            const uint32_t dg = P.vdeg[i];
            const uint32_t id = (uint32_t)lane < dg ? P.vlinks[(size_t)i * P.vM + lane] : U32_NONE;
            if (visit(id, id != U32_NONE && id < P.n && id != i, 0.0f)) process(false);
        }
    }
    if (status) {
        if (lane == 0) { P.status[i] = (uint32_t)status; P.out_cnt[i] = 0; }
        return;
    }
    if (P.work && lane == 0) {
        atomicAdd(P.work, (unsigned long long)nd2);
        atomicAdd(P.work + 1, (unsigned long long)(nadj + P.cap));   // the seeds: the row's own slots
    }
    if (!GRAPH) {
        if constexpr (PCH > 0) {
#pragma unroll
            for (int j = 0; j < PCH; j++) {
                const uint32_t t = 64u * j + (uint32_t)lane;
                if (t < P.K) {
                    P.out_ids[(size_t)i * P.K + t] = t < psize ? pkey_id(PL.k[j]) : U32_NONE;
                    P.out_keys[(size_t)i * P.K + t] = t < psize ? pkey_d(PL.k[j]) : F32_BIG;
                }
            }
        } else {
            const float* ad = pb ? S.pd2 : S.pd;
            const uint32_t* ap = pb ? S.pi2 : S.pi;
            for (uint32_t t = lane; t < P.K; t += 64) {
                P.out_ids[(size_t)i * P.K + t] = t < psize ? ap[t] : U32_NONE;
                P.out_keys[(size_t)i * P.K + t] = t < psize ? ad[t] : F32_BIG;
            }
        }
        if (lane == 0) { P.out_cnt[i] = psize; P.status[i] = 0; }
        return;
    }
    // graph mode: drop self, sort by (dist, id), alpha-RNG to M
    uint32_t w = 0;
    for (uint32_t t0 = 0; t0 < psize; t0 += 64) {
        uint32_t t = t0 + lane;
        int keep = t < psize && S.pi[t] != i;
        float dv = keep ? S.pd[t] : 0.0f;
        uint32_t id = keep ? S.pi[t] : 0u;
        uint64_t km = ballot(keep);
        __syncthreads();
        if (keep) { S.pd[w + lanes_below(km)] = dv; S.pi[w + lanes_below(km)] = id; }
        w += (uint32_t)popc64(km);
        __syncthreads();
    }
    uint32_t cap2 = 64;
    while (cap2 < w) cap2 <<= 1;
    if (cap2 > P.pmax) { if (lane == 0) { P.status[i] = 2; P.out_cnt[i] = 0; } return; }
    lds_sort_pairs(S.pd, S.pi, w, cap2);
    uint32_t deg;
    if (P.pol.on) {
        // the work area ends with the policy's scratch (global work area only)
        __shared__ float pol_nrm[64];
        deg = policy_select(P.x, P.nrm, P.n, d, P.M, P.alpha, P.pol, i, S.pd, S.pi, w, S.pscratch, cap2, S.selv, S.sel,
                            pol_nrm, P.out_ids + (size_t)i * P.M);
    } else {
        deg = rng_select<MC>(P, S.pd, S.pi, w, S.selv, S.sel, P.out_ids + (size_t)i * P.M);
    }
    if (lane == 0) { P.out_cnt[i] = deg; P.status[i] = 0; }
}

template <bool GRAPH, bool GWS, int MC, int QCH, int PCH, int HM>
__global__ void __launch_bounds__(64) k_gather(GatherParams P) {
    extern __shared__ __align__(16) char lds_raw[];
    uint32_t bi = blockIdx.x;
    if (bi >= P.nrows) return;
    uint32_t i = P.rows ? P.rows[bi] : bi;
    char* base = GWS ? (char*)(P.gscratch + (size_t)bi * P.gscratch_words) : lds_raw;
    GatherLds S = gather_carve(base, P, GRAPH, QCH > 0);
    gather_row<GRAPH, GWS, MC, QCH, PCH, HM>(P, S, i);
}

template <bool GRAPH, bool GWS, int MC, int QCH, int PCH, int HM = 0>
static hipError_t launch_gather_t(const GatherParams& P, size_t lds, hipStream_t s) {
    const void* f = (const void*)k_gather<GRAPH, GWS, MC, QCH, PCH, HM>;
    if (lds) {
        if (hipError_t e = set_lds_limit(f, lds); e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_gather<GRAPH, GWS, MC, QCH, PCH, HM>), dim3(P.nrows), dim3(64), lds, s, P);
    return hipGetLastError();
}

template <bool GRAPH, bool GWS, int MC>
static hipError_t launch_gather_lists(const GatherParams& P, size_t lds, hipStream_t s) {
    if constexpr (!GWS) {
        if (gather_rk(P, GRAPH)) {
            constexpr int PC = GRAPH ? 0 : 2;
            if (!GRAPH && P.h16 == 1) {
                if (P.explore <= 128) return launch_gather_t<GRAPH, GWS, MC, 2, PC, 1>(P, lds, s);
                return launch_gather_t<GRAPH, GWS, MC, 10, PC, 1>(P, lds, s);
            }
            if (!GRAPH && P.h16 == 2) {
                if (P.explore <= 128) return launch_gather_t<GRAPH, GWS, MC, 2, PC, 2>(P, lds, s);
                return launch_gather_t<GRAPH, GWS, MC, 10, PC, 2>(P, lds, s);
            }
            if (P.explore <= 128) return launch_gather_t<GRAPH, GWS, MC, 2, PC>(P, lds, s);
            return launch_gather_t<GRAPH, GWS, MC, 10, PC>(P, lds, s);
        }
    }
    return launch_gather_t<GRAPH, GWS, MC, 0, 0>(P, lds, s);
}

template <bool GRAPH, bool GWS>
static hipError_t launch_gather_mc(const GatherParams& P, size_t lds, hipStream_t s) {
    switch (P.d >> 4) {
        case 8: return launch_gather_lists<GRAPH, GWS, 8>(P, lds, s);
        case 6: return launch_gather_lists<GRAPH, GWS, 6>(P, lds, s);
        default: return launch_gather_lists<GRAPH, GWS, 0>(P, lds, s);
    }
}

hipError_t launch_gather(const GatherParams& P, bool graph, bool gws, hipStream_t s) {
    if (P.nrows == 0) return hipSuccess;
    size_t lds = gws ? 0 : gather_lds_bytes(P, graph, gather_rk(P, graph));
    if (graph) return gws ? launch_gather_mc<true, true>(P, lds, s) : launch_gather_mc<true, false>(P, lds, s);
    return gws ? launch_gather_mc<false, true>(P, lds, s) : launch_gather_mc<false, false>(P, lds, s);
}

// ---- commit: cand = O_i u R_i, prune to M (include/index.h:907-930) -----------------
size_t commit_lds_bytes(const PruneParams& P) {
    return 2 * rnd16(4 * (size_t)P.pmax) + rnd16(4 * (size_t)P.M * P.d) + rnd16(4 * (size_t)P.M) +
           rnd16(4 * 64) + (P.pol.on ? rnd16(policy_scratch_bytes(std::max<uint32_t>(P.pmax, 64), P.pol, P.M)) : 0);
}

// MC = (d >> 4) for the batched distance pass (0: one row per group at a time).
template <bool GWS, int MC>
__global__ void __launch_bounds__(64) k_commit(PruneParams P) {
    extern __shared__ __align__(16) char lds_raw[];
    uint32_t bi = blockIdx.x;
    if (bi >= P.nrows) return;
    uint32_t i = P.rows ? P.rows[bi] : bi;
    char* base = GWS ? (char*)(P.gscratch + (size_t)bi * P.gscratch_words) : lds_raw;
    float* pd = (float*)base;
    uint32_t* pi = (uint32_t*)(base + rnd16(4 * (size_t)P.pmax));
    float* selv = (float*)(base + 2 * rnd16(4 * (size_t)P.pmax));
    uint32_t* sel = (uint32_t*)((char*)selv + rnd16(4 * (size_t)P.M * P.d));
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    const int d = (int)P.d;
    uint32_t no = P.ocnt[i], r0 = P.roff[i], r1 = P.roff[i + 1];
    uint32_t m = no + (r1 - r0);
    uint32_t cap2 = 64;
    while (cap2 < m) cap2 <<= 1;
    // on overflow the row keeps its degree: O may be the links themselves
    if (cap2 > P.pmax) { if (lane == 0) P.status[i] = 2; return; }
    for (uint32_t t = lane; t < m; t += 64) {
        const uint32_t r = t < no ? 0u : P.rids[r0 + (t - no)];
        pi[t] = t < no ? P.O[(size_t)i * P.M + t] : (P.rmap ? P.rmap[r] : r);
    }
    __syncthreads();
    QRow QI;
    qrow_load(QI, P.x + (size_t)i * d, d, l);
    float ni = P.nrm[i];
    if constexpr (MC > 0) {
        // four rows per group in flight (same FMA order and tree as qrow_ip)
        rows_batch<MC, 1>(QI, P.x, (uint32_t)d, P.nrm, ni, pi, (int)m, l, g, pd);
    } else {
        for (uint32_t s = g; s < m; s += 4) {
            uint32_t j = pi[s];
            float ip = qrow_ip(QI, P.x + (size_t)j * d, d, l);
            if (l == 0) pd[s] = d2_combine(ip, ni, P.nrm[j]);
        }
    }
    __syncthreads();
    lds_sort_pairs(pd, pi, m, cap2);
    // drop self and duplicates (same id -> same distance -> adjacent)
    uint32_t w = 0;
    for (uint32_t t0 = 0; t0 < m; t0 += 64) {
        uint32_t t = t0 + lane;
        int keep = t < m && pi[t] != i && pi[t] < P.n && !(t > 0 && pi[t - 1] == pi[t]);
        float dv = keep ? pd[t] : 0.0f;
        uint32_t id = keep ? pi[t] : 0u;
        uint64_t km = ballot(keep);
        __syncthreads();
        if (keep) { pd[w + lanes_below(km)] = dv; pi[w + lanes_below(km)] = id; }
        w += (uint32_t)popc64(km);
        __syncthreads();
    }
    uint32_t dg;
    if (P.pol.on) {
        __shared__ float pol_nrm[64];
        char* scratch = base + 2 * rnd16(4 * (size_t)P.pmax) + rnd16(4 * (size_t)P.M * P.d) + rnd16(4 * (size_t)P.M) +
                        rnd16(4 * 64);   // the last region of commit_lds_bytes
        dg = policy_select(P.x, P.nrm, P.n, d, P.M, P.alpha, P.pol, i, pd, pi, w, scratch, cap2, selv, sel, pol_nrm,
                           P.links + (size_t)i * P.M);
    } else {
        GatherParams G{};
        G.x = P.x; G.nrm = P.nrm; G.n = P.n; G.d = P.d; G.M = P.M; G.alpha = P.alpha;
        dg = rng_select<MC>(G, pd, pi, w, selv, sel, P.links + (size_t)i * P.M);
    }
    for (uint32_t t = dg + (uint32_t)lane; t < P.M; t += 64) P.links[(size_t)i * P.M + t] = U32_NONE;
    if (lane == 0) { P.deg[i] = dg; P.status[i] = 0; }
}

template <int MC>
static hipError_t launch_commit_mc(const PruneParams& P, bool gws, hipStream_t s) {
    size_t lds = gws ? 0 : commit_lds_bytes(P);
    if (gws) {
        hipLaunchKernelGGL((k_commit<true, MC>), dim3(P.nrows), dim3(64), 0, s, P);
    } else {
        if (hipError_t e = set_lds_limit((const void*)k_commit<false, MC>, lds); e != hipSuccess) return e;
        hipLaunchKernelGGL((k_commit<false, MC>), dim3(P.nrows), dim3(64), lds, s, P);
    }
    return hipGetLastError();
}

hipError_t launch_commit(const PruneParams& P, bool gws, hipStream_t s) {
    if (P.nrows == 0) return hipSuccess;
    switch (P.d >> 4) {
        case 8: return launch_commit_mc<8>(P, gws, s);
        case 6: return launch_commit_mc<6>(P, gws, s);
        default: return launch_commit_mc<0>(P, gws, s);
    }
}

// ---- Vamana refinement helpers (include/index.h:190-222, 831-1065) ---------------------

// The search-tree entry of each row (rptrees.h:108-123, ties go left), the
// row itself descending: four rows per wavefront, one per 16-lane group.
// MC = dl >> 4 (or 0): each level's loads are issued together.
template <int MC>
__global__ void __launch_bounds__(64) k_tree_entry_rows(const float* __restrict__ x, uint32_t n, uint32_t dl,
                                                        const uint32_t* __restrict__ left,
                                                        const uint32_t* __restrict__ right,
                                                        const float* __restrict__ off,
                                                        const float* __restrict__ hyper,
                                                        const uint32_t* __restrict__ entry, uint32_t root,
                                                        uint32_t* __restrict__ out) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    const uint32_t i = blockIdx.x * 4u + (uint32_t)g;
    if (i >= n) return;                         // whole 16-lane groups
    QRow QL;
    qrow_load_mc<MC>(QL, x + (size_t)i * dl, (int)dl, l);
    uint32_t node = root;
    for (;;) {
        const uint32_t lc = left[node], rc = right[node];
        const float ip = qrow_ip_mc<MC>(QL, hyper + (size_t)node * dl, (int)dl, l);
        if (lc == U32_NONE) break;
        const float mg = ip - off[node];
        node = (mg < -1e-6f) ? lc : (mg > 1e-6f ? rc : lc);
    }
    if (l == 0) out[i] = entry[node];
}

hipError_t launch_tree_entry_rows(const float* x, uint32_t n, uint32_t dl, const uint32_t* left,
                                  const uint32_t* right, const float* off, const float* hyper,
                                  const uint32_t* entry, uint32_t root, uint32_t* out, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 3) / 4);
    if ((dl >> 4) == 8)
        hipLaunchKernelGGL(k_tree_entry_rows<8>, grid, dim3(64), 0, s, x, n, dl, left, right, off, hyper, entry, root, out);
    else if ((dl >> 4) == 6)
        hipLaunchKernelGGL(k_tree_entry_rows<6>, grid, dim3(64), 0, s, x, n, dl, left, right, off, hyper, entry, root, out);
    else
        hipLaunchKernelGGL(k_tree_entry_rows<0>, grid, dim3(64), 0, s, x, n, dl, left, right, off, hyper, entry, root, out);
    return hipGetLastError();
}

// write_ids (index.h:897-903) for a batch, one wavefront per row.
__global__ void __launch_bounds__(64) k_vamana_write(const uint32_t* __restrict__ rows, uint32_t nb,
                                                     const uint32_t* __restrict__ O,
                                                     const uint32_t* __restrict__ ocnt, uint32_t M,
                                                     uint32_t* __restrict__ links, uint32_t* __restrict__ deg,
                                                     uint32_t* __restrict__ Ob, uint32_t* __restrict__ obcnt) {
    const uint32_t b = blockIdx.x, lane = (uint32_t)__lane_id();
    if (b >= nb) return;
    const uint32_t node = rows[b], c = ocnt[node];
    for (uint32_t t = lane; t < M; t += 64) {
        const uint32_t v = t < c ? O[(size_t)node * M + t] : U32_NONE;
        links[(size_t)node * M + t] = v;
        Ob[(size_t)b * M + t] = v;
    }
    if (lane == 0) { deg[node] = c; obcnt[b] = c; }
}

hipError_t launch_vamana_write(const uint32_t* rows, uint32_t nb, const uint32_t* O, const uint32_t* ocnt,
                               uint32_t M, uint32_t* links, uint32_t* deg, uint32_t* Ob, uint32_t* obcnt,
                               hipStream_t s) {
    if (nb == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vamana_write, dim3(nb), dim3(64), 0, s, rows, nb, O, ocnt, M, links, deg, Ob, obcnt);
    return hipGetLastError();
}

__global__ void __launch_bounds__(256) k_nonempty_rows(const uint32_t* __restrict__ roff, uint32_t n,
                                                       uint32_t* __restrict__ list, uint32_t* __restrict__ cnt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const int has = i < n && roff[i + 1] > roff[i];
    const uint64_t m = ballot(has);
    uint32_t base = 0;
    if (__lane_id() == 0 && m) base = atomicAdd(cnt, (uint32_t)popc64(m));
    base = __shfl(base, 0);
    if (has) list[base + lanes_below(m)] = i;
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

hipError_t launch_nonempty_rows(const uint32_t* roff, uint32_t n, uint32_t* list, uint32_t* cnt, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, cnt);
    if (n) hipLaunchKernelGGL(k_nonempty_rows, dim3((n + 255) / 256), dim3(256), 0, s, roff, n, list, cnt);
    return hipGetLastError();
}

// choose_vamana_entry's medoid: sampled rows (t step) % n.  The mean of each
// coordinate is summed in sample order in binary64, one thread per
// coordinate; each sampled row's squared distance to it is a fused binary64
// sum in coordinate order (the reference's loop as GCC contracts it).
__global__ void __launch_bounds__(64) k_vamana_mean(const float* __restrict__ x, uint32_t n, uint32_t dl,
                                                    uint64_t sample, uint64_t step, double* __restrict__ mean) {
    const uint32_t j = blockIdx.x * 64u + (uint32_t)__lane_id();
    if (j >= dl) return;
    double acc = 0.0;
    for (uint64_t t = 0; t < sample; t++) acc += (double)x[((t * step) % n) * dl + j];
    mean[j] = acc / (double)sample;
}

__device__ __forceinline__ double vamana_dd(const float* r, const double* mean, uint32_t dl) {
    double dd = 0.0;
    for (uint32_t j = 0; j < dl; j++) {
        const double z = (double)r[j] - mean[j];
        dd = fma(z, z, dd);
    }
    return dd;
}

// pass 0: best[0] = least distance (bit patterns of non-negative doubles
// order as integers); pass 1: best[1] = least t attaining it.
__global__ void __launch_bounds__(256) k_vamana_medoid(const float* __restrict__ x, uint32_t n, uint32_t dl,
                                                       uint64_t sample, uint64_t step,
                                                       const double* __restrict__ mean,
                                                       unsigned long long* __restrict__ best, int pass) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= sample) return;
    const double dd = vamana_dd(x + ((t * step) % n) * dl, mean, dl);
    const unsigned long long b = (unsigned long long)__double_as_longlong(dd);
    if (pass == 0) atomicMin(best, b);
    else if (b == best[0]) atomicMin(best + 1, (unsigned long long)t);
}

__global__ void k_vamana_medoid_init(unsigned long long* best) { best[0] = ~0ull; best[1] = ~0ull; }

hipError_t launch_vamana_medoid(const float* x, uint32_t n, uint32_t dl, uint64_t sample, double* mean,
                                unsigned long long* best, uint32_t* out, hipStream_t s) {
    if (n == 0) return hipErrorInvalidValue;
    if (sample < 1) sample = 1;
    if (sample > n) sample = n;
    const uint64_t step = n / sample >= 1 ? n / sample : 1;
    hipLaunchKernelGGL(k_vamana_mean, dim3((dl + 63) / 64), dim3(64), 0, s, x, n, dl, sample, step, mean);
    hipLaunchKernelGGL(k_vamana_medoid_init, dim3(1), dim3(1), 0, s, best);
    const dim3 grid((uint32_t)((sample + 255) / 256));
    hipLaunchKernelGGL(k_vamana_medoid, grid, dim3(256), 0, s, x, n, dl, sample, step, mean, best, 0);
    hipLaunchKernelGGL(k_vamana_medoid, grid, dim3(256), 0, s, x, n, dl, sample, step, mean, best, 1);
    unsigned long long bt = 0;
    if (hipError_t e = hipMemcpyAsync(&bt, best + 1, 8, hipMemcpyDeviceToHost, s); e != hipSuccess) return e;
    if (hipError_t e = hipStreamSynchronize(s); e != hipSuccess) return e;
    *out = (uint32_t)((bt * step) % n);
    return hipGetLastError();
}
