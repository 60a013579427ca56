// k_forest.hip -- random-projection forest (include/rptrees.h:27-184) built
// level-synchronously for all trees at once, and the leaf all-pairs kernel of
// init_knn (include/index.h:1433-1460).
//
// Split rule (rptrees.h:27-84, L2): two distinct random members o0, o1 of the
// node; normal h = o0 - o1, offset = <h, (o0 + o1) / 2>; margin = <h, x> -
// offset; margin < -EPS left, > EPS right, otherwise a coin flip; if one side
// ends up empty every member is re-dealt by coin flips.  Partitions are
// stable, so each leaf keeps its members in ascending original order (the
// reference's indices order).  Randomness comes from a counter-based
// generator keyed by (seed, tree, node), not the reference's thread-local
// mt19937, so forests agree with the reference in distribution only.
//
// Leaf kernel: for every row of a leaf, exact FP32 distances to all other
// members (the reference's d2), candidates that beat the row's current K-th
// (dist, id) pair are sorted and merged into the row's sorted top-K list.
// A pair seen in an earlier tree has identical distance bits, so duplicates
// are adjacent after sorting and dropped.  Rows are owned by exactly one leaf
// per tree, so one launch per tree needs no atomics.
#include "common.hpp"
#include "kernels.hpp"

// The "ip" metric's split (rptrees.h:41-52): h = o0 / |o0| - o1 / |o1|,
// normalised, offset 0 (norms below EPS taken as 1).  Lane l of the group
// holds elements l, 16 + l, ... of h in registers; its squared norm is
// reduced with the inner-product tree.
__device__ void angular_plane(const ForestLevel& F, const float* a, const float* b, float* h, uint32_t sgi,
                              int l) {
    const int d = (int)F.d, m = d >> 4, r = d & 15;
    float n0 = sqrtf(g16_ip(a, a, d, l));
    float n1 = sqrtf(g16_ip(b, b, d, l));
    if (fabsf(n0) < 1e-6f) n0 = 1.0f;
    if (fabsf(n1) < 1e-6f) n1 = 1.0f;
    float hv[QMAX_CHUNKS + 1];
    float acc = 0.0f;
#pragma unroll
    for (int t = 0; t < QMAX_CHUNKS; t++) {
        hv[t] = 0.0f;
        if (t < m) {
            hv[t] = a[16 * t + l] / n0 - b[16 * t + l] / n1;
            acc = fmaf(hv[t], hv[t], acc);
        }
    }
    float hn2 = tree16(acc);
    hv[QMAX_CHUNKS] = (l < r) ? a[16 * m + l] / n0 - b[16 * m + l] / n1 : 0.0f;
    if (r) hn2 += tree16(hv[QMAX_CHUNKS] * hv[QMAX_CHUNKS]);
    float hn = sqrtf(hn2);
    if (fabsf(hn) < 1e-6f) hn = 1.0f;
#pragma unroll
    for (int t = 0; t < QMAX_CHUNKS; t++)
        if (t < m) h[16 * t + l] = hv[t] / hn;
    if (l < r) h[16 * m + l] = hv[QMAX_CHUNKS] / hn;
    if (l == 0) F.seg_off[sgi] = 0.0f;
}

// One 16-lane group per segment: hyperplane and offset.
__global__ void __launch_bounds__(256) k_split_planes(ForestLevel F) {
    uint32_t sgi = (blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    int l = threadIdx.x & 15;
    if (sgi >= F.nseg) return;
    uint32_t start = F.seg_start[sgi];
    uint32_t p0 = F.perm[start + F.seg_pick[2 * sgi]];
    uint32_t p1 = F.perm[start + F.seg_pick[2 * sgi + 1]];
    const float* a = F.x + (size_t)p0 * F.d;
    const float* b = F.x + (size_t)p1 * F.d;
    float* h = F.seg_hyper + (size_t)sgi * F.d;
    if (F.angular) {
        angular_plane(F, a, b, h, sgi, l);
        return;
    }
    for (uint32_t k = l; k < F.d; k += 16) h[k] = a[k] - b[k];
    // offset = <h, midpoint>, midpoint = (a + b) / 2 (exact halving)
    int m = (int)F.d >> 4, r = (int)F.d & 15;
    float acc = 0.0f;
    for (int t = 0; t < m; t++) {
        float mid = (a[16 * t + l] + b[16 * t + l]) / 2.0f;
        acc = fmaf(a[16 * t + l] - b[16 * t + l], mid, acc);
    }
    float res = tree16(acc);
    if (r) {
        float p = 0.0f;
        if (l < r) {
            float mid = (a[16 * m + l] + b[16 * m + l]) / 2.0f;
            p = (a[16 * m + l] - b[16 * m + l]) * mid;
        }
        res += tree16(p);
    }
    if (l == 0) F.seg_off[sgi] = res;
}

// One 16-lane group per position of an active segment: side 0 (left) / 1.
__global__ void __launch_bounds__(256) k_split_sides(ForestLevel F, uint64_t seed) {
    uint64_t pos = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 4;
    int l = threadIdx.x & 15;
    if (pos >= F.total) return;
    uint32_t sgi = F.seg_of[pos];
    if (sgi == U32_NONE) return;
    uint32_t id = F.perm[pos];
    float mg = g16_ip(F.seg_hyper + (size_t)sgi * F.d, F.x + (size_t)id * F.d, (int)F.d, l) - F.seg_off[sgi];
    uint8_t side;
    if (mg < -1e-6f) side = 0;
    else if (mg > 1e-6f) side = 1;
    else side = (uint8_t)(rng_at(seed, ((uint64_t)sgi << 1) | 1, pos) & 1);
    if (l == 0) F.side[pos] = side;
}

hipError_t launch_split_planes(const ForestLevel& F, hipStream_t s) {
    if (F.nseg == 0) return hipSuccess;
    hipLaunchKernelGGL(k_split_planes, dim3((unsigned)(((uint64_t)F.nseg * 16 + 255) / 256)), dim3(256), 0, s, F);
    return hipGetLastError();
}
hipError_t launch_split_sides(const ForestLevel& F, uint64_t seed, hipStream_t s) {
    if (F.total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_split_sides, dim3((unsigned)((F.total * 16 + 255) / 256)), dim3(256), 0, s, F, seed);
    return hipGetLastError();
}

// ---- leaf all-pairs top-K ------------------------------------------------------------
__host__ __device__ __forceinline__ size_t lrnd(size_t b) { return (b + 15) & ~size_t(15); }

size_t leaf_lds_bytes(const LeafParams& P) {
    uint32_t cap2 = 64;
    while (cap2 < P.max_leaf) cap2 <<= 1;
    // per wave: candidate buffer (cap2 pairs) + list double buffer (2 x K pairs)
    return 4 * (lrnd(8 * (size_t)cap2) + 2 * lrnd(8 * (size_t)P.K));
}

__device__ void wave_sort_lds(float* pd, uint32_t* pi, uint32_t m, uint32_t cap2) {
    const int lane = __lane_id();
    for (uint32_t t = m + lane; t < cap2; t += 64) { pd[t] = F32_BIG; pi[t] = U32_NONE; }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
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
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        }
    }
}

#define WAVE_SYNC()                                              \
    do {                                                         \
        __builtin_amdgcn_wave_barrier();                         \
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");   \
    } while (0)

// 4 wavefronts per leaf; wavefront w handles rows w, w+4, ...  MC = d >> 4
// (or 0): a member's row is requested whole before its dot product starts.
template <int MC>
__global__ void __launch_bounds__(256) k_leaf_knn(LeafParams P) {
    extern __shared__ __align__(16) char lds[];
    const uint32_t leaf = blockIdx.x;
    if (leaf >= P.nleaves) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, l = lane & 15, g = lane >> 4;
    uint32_t cap2 = 64;
    while (cap2 < P.max_leaf) cap2 <<= 1;
    size_t per = lrnd(8 * (size_t)cap2) + 2 * lrnd(8 * (size_t)P.K);
    char* mine = lds + per * w;
    float* cd = (float*)mine;
    uint32_t* ci = (uint32_t*)(mine + 4 * (size_t)cap2);
    float* ld0 = (float*)(mine + lrnd(8 * (size_t)cap2));
    uint32_t* li0 = (uint32_t*)((char*)ld0 + 4 * (size_t)P.K);
    float* ld1 = (float*)((char*)ld0 + lrnd(8 * (size_t)P.K));
    uint32_t* li1 = (uint32_t*)((char*)ld1 + 4 * (size_t)P.K);
    const uint32_t start = P.leaf_start[leaf], size = P.leaf_size[leaf];
    const uint32_t* mem = P.perm + start;
    const int d = (int)P.d;
    for (uint32_t a = (uint32_t)w; a < size; a += 4) {
        uint32_t r = mem[a];
        uint32_t cnt = P.cnt[r];
        uint32_t* rid = P.ids + (size_t)r * P.K;
        float* rkd = P.keys + (size_t)r * P.K;
        uint64_t thr = cnt >= P.K ? pkey(rkd[P.K - 1], rid[P.K - 1]) : ~0ull;
        QRow QR;
        qrow_load_mc<MC>(QR, P.x + (size_t)r * d, d, l);
        float nr = P.nrm[r];
        uint32_t nc = 0;
        if (P.ipm) {
            // "ip": one member per lane, 1 - <x_r, x_c> summed in element order
            for (uint32_t b0 = 0; b0 < size; b0 += 64) {
                const uint32_t b = b0 + (uint32_t)lane;
                int ok = 0;
                float dv = 0.0f;
                uint32_t c = 0;
                if (b < size && b != a) {
                    c = mem[b];
                    dv = ipdist_seq(P.x + (size_t)r * d, P.x + (size_t)c * d, (uint32_t)d);
                    ok = pkey(dv, c) < thr;
                }
                uint64_t m = ballot(ok);
                if (ok) { uint32_t p = nc + lanes_below(m); cd[p] = dv; ci[p] = c; }
                nc += (uint32_t)popc64(m);
            }
        } else for (uint32_t b0 = 0; b0 < size; b0 += 4) {
            uint32_t b = b0 + (uint32_t)g;
            int ok = 0;
            float dv = 0.0f;
            uint32_t c = 0;
            if (b < size && b != a) {
                c = mem[b];
                float ip = qrow_ip_mc<MC>(QR, P.x + (size_t)c * d, d, l);
                dv = d2_combine(ip, nr, P.nrm[c]);
                ok = (l == 0) && (pkey(dv, c) < thr);
            }
            uint64_t m = ballot(ok);
            if (ok) { uint32_t p = nc + lanes_below(m); cd[p] = dv; ci[p] = c; }
            nc += (uint32_t)popc64(m);
        }
        WAVE_SYNC();
        if (nc == 0) continue;
        uint32_t c2 = 64;
        while (c2 < nc) c2 <<= 1;
        wave_sort_lds(cd, ci, nc, c2);
        // load the row list
        for (uint32_t t = lane; t < cnt; t += 64) { ld0[t] = rkd[t]; li0[t] = rid[t]; }
        WAVE_SYNC();
        // drop candidates already present (identical pair -> identical bits)
        uint32_t kept = 0;
        for (uint32_t t0 = 0; t0 < nc; t0 += 64) {
            uint32_t t = t0 + lane;
            int keep = 0;
            float dv = 0.0f;
            uint32_t id = 0;
            if (t < nc) {
                dv = cd[t]; id = ci[t];
                uint64_t key = pkey(dv, id);
                uint32_t lo = 0, hi = cnt;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (pkey(ld0[mid], li0[mid]) < key) lo = mid + 1; else hi = mid;
                }
                keep = !(lo < cnt && li0[lo] == id);
            }
            uint64_t km = ballot(keep);
            WAVE_SYNC();
            if (keep) { cd[kept + lanes_below(km)] = dv; ci[kept + lanes_below(km)] = id; }
            kept += (uint32_t)popc64(km);
            WAVE_SYNC();
        }
        if (kept == 0) continue;
        // merge (cd, ci)[0..kept) with the list into (ld1, li1)
        for (uint32_t t = lane; t < cnt; t += 64) {
            uint64_t bk = pkey(ld0[t], li0[t]);
            uint32_t lo = 0, hi = kept;
            while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (pkey(cd[mid], ci[mid]) < bk) lo = mid + 1; else hi = mid; }
            uint32_t pos = t + lo;
            if (pos < P.K) { ld1[pos] = ld0[t]; li1[pos] = li0[t]; }
        }
        for (uint32_t t = lane; t < kept; t += 64) {
            uint64_t ck = pkey(cd[t], ci[t]);
            uint32_t lo = 0, hi = cnt;
            while (lo < hi) { uint32_t mid = (lo + hi) >> 1; if (pkey(ld0[mid], li0[mid]) < ck) lo = mid + 1; else hi = mid; }
            uint32_t pos = t + lo;
            if (pos < P.K) { ld1[pos] = cd[t]; li1[pos] = ci[t]; }
        }
        WAVE_SYNC();
        uint32_t ncnt = min(P.K, cnt + kept);
        for (uint32_t t = lane; t < ncnt; t += 64) { rkd[t] = ld1[t]; rid[t] = li1[t]; }
        if (lane == 0) P.cnt[r] = ncnt;
        WAVE_SYNC();
    }
}

hipError_t launch_leaf_knn(const LeafParams& P, bool mfma, hipStream_t s) {
    (void)mfma;
    if (P.nleaves == 0) return hipSuccess;
    size_t lds = leaf_lds_bytes(P);
    const uint32_t mc = P.d >> 4;
    const void* f = mc == 8 ? (const void*)k_leaf_knn<8> : mc == 6 ? (const void*)k_leaf_knn<6>
                                                                   : (const void*)k_leaf_knn<0>;
    if (hipError_t e = set_lds_limit(f, lds); e != hipSuccess) return e;
    if (mc == 8) hipLaunchKernelGGL(k_leaf_knn<8>, dim3(P.nleaves), dim3(256), lds, s, P);
    else if (mc == 6) hipLaunchKernelGGL(k_leaf_knn<6>, dim3(P.nleaves), dim3(256), lds, s, P);
    else hipLaunchKernelGGL(k_leaf_knn<0>, dim3(P.nleaves), dim3(256), lds, s, P);
    return hipGetLastError();
}

// init_knn's fill for short rows (index.h:1452-1460): K - cnt random draws,
// each inserted unless already present.  Rare; one thread per row.
__global__ void k_random_fill(const float* __restrict__ x, const float* __restrict__ nrm, uint32_t n,
                              uint32_t d, uint32_t K, uint64_t seed, int ipm, uint32_t* __restrict__ ids,
                              float* __restrict__ keys, uint32_t* __restrict__ cnt) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c = cnt[i];
    if (c >= K) return;
    uint32_t miss = K - c;
    uint32_t* rid = ids + (size_t)i * K;
    float* rk = keys + (size_t)i * K;
    for (uint32_t j = 0; j < miss; j++) {
        uint32_t r = (uint32_t)(rng_at(seed, i, j) % n);
        bool dup = false;
        for (uint32_t t = 0; t < c; t++) dup |= (rid[t] == r);
        if (dup) continue;
        const float* a = x + (size_t)i * d;
        const float* b = x + (size_t)r * d;
        float dv;
        if (ipm) {
            dv = ipdist_seq(a, b, d);
        } else {
        // sequential FP32 contract (16 lanes emulated in registers)
        float acc[16];
        for (int q = 0; q < 16; q++) acc[q] = 0.0f;
        uint32_t m = d >> 4, rr = d & 15;
        for (uint32_t t = 0; t < m; t++)
            for (int q = 0; q < 16; q++) acc[q] = fmaf(a[16 * t + q], b[16 * t + q], acc[q]);
        float t3[8], t6[4];
        for (int q = 0; q < 8; q++) t3[q] = acc[q + 8] + acc[q];
        for (int q = 0; q < 4; q++) t6[q] = t3[q + 4] + t3[q];
        float ip = (t6[0] + t6[2]) + (t6[1] + t6[3]);
        if (rr) {
            float p[16];
            for (int q = 0; q < 16; q++) p[q] = (uint32_t)q < rr ? a[16 * m + q] * b[16 * m + q] : 0.0f;
            for (int q = 0; q < 8; q++) t3[q] = p[q + 8] + p[q];
            for (int q = 0; q < 4; q++) t6[q] = t3[q + 4] + t3[q];
            ip += (t6[0] + t6[2]) + (t6[1] + t6[3]);
        }
        dv = fmaf(-2.0f, ip, nrm[i] + nrm[r]);
        }
        // sorted insertion
        uint32_t p = c;
        while (p > 0 && (rk[p - 1] > dv || (rk[p - 1] == dv && rid[p - 1] > r))) {
            rk[p] = rk[p - 1]; rid[p] = rid[p - 1]; p--;
        }
        rk[p] = dv; rid[p] = r;
        c++;
    }
    cnt[i] = c;
}

hipError_t launch_random_fill(const float* x, const float* nrm, uint32_t n, uint32_t d, uint32_t K,
                              uint64_t seed, int ipm, uint32_t* ids, float* keys, uint32_t* cnt, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_random_fill, dim3((n + 127) / 128), dim3(128), 0, s, x, nrm, n, d, K, seed, ipm, ids,
                       keys, cnt);
    return hipGetLastError();
}

// ---- level bookkeeping ---------------------------------------------------------------
__global__ void k_seg_fill(uint32_t* __restrict__ seg_of, const uint32_t* __restrict__ seg_start,
                           const uint32_t* __restrict__ seg_size, uint32_t nseg) {
    uint32_t s = blockIdx.x;
    if (s >= nseg) return;
    uint32_t st = seg_start[s], sz = seg_size[s];
    for (uint32_t t = threadIdx.x; t < sz; t += blockDim.x) seg_of[(uint64_t)st + t] = s;
}

// 16 consecutive positions per thread, one atomic per run of equal segments.
__global__ void k_seg_count_left(const uint8_t* __restrict__ side, const uint32_t* __restrict__ seg_of,
                                 uint64_t total, uint32_t* __restrict__ left) {
    uint64_t p0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
    uint32_t cur = U32_NONE, acc = 0;
    for (uint64_t p = p0; p < p0 + 16 && p < total; p++) {
        uint32_t s = seg_of[p];
        if (s != cur) {
            if (cur != U32_NONE && acc) atomicAdd(&left[cur], acc);
            cur = s;
            acc = 0;
        }
        if (s != U32_NONE && side[p] == 0) acc++;
    }
    if (cur != U32_NONE && acc) atomicAdd(&left[cur], acc);
}

// Re-deal every member of a degenerate segment by coin flips (rptrees.h:76-81).
__global__ void k_seg_redeal(uint8_t* __restrict__ side, const uint32_t* __restrict__ seg_of,
                             const uint8_t* __restrict__ degen, uint64_t total, uint64_t seed) {
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    uint32_t s = seg_of[p];
    if (s == U32_NONE || !degen[s]) return;
    side[p] = (uint8_t)(rng_at(seed, ((uint64_t)s << 1), p) & 1);
}

__global__ void k_left_flags(const uint8_t* __restrict__ side, const uint32_t* __restrict__ seg_of,
                             uint64_t total, uint32_t* __restrict__ flags) {
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    flags[p] = (seg_of[p] != U32_NONE && side[p] == 0) ? 1u : 0u;
}

// Stable partition of every active segment (lscan: exclusive scan of flags).
__global__ void k_seg_partition(const uint32_t* __restrict__ perm, uint32_t* __restrict__ perm_out,
                                const uint8_t* __restrict__ side, const uint32_t* __restrict__ seg_of,
                                const uint32_t* __restrict__ seg_start, const uint32_t* __restrict__ left,
                                const uint32_t* __restrict__ lscan, uint64_t total) {
    uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= total) return;
    uint32_t s = seg_of[p];
    if (s == U32_NONE) { perm_out[p] = perm[p]; return; }
    uint64_t st = seg_start[s];
    uint32_t lr = lscan[p] - lscan[st];
    uint64_t np = side[p] == 0 ? st + lr : st + left[s] + ((p - st) - lr);
    perm_out[np] = perm[p];
}

hipError_t launch_seg_fill(uint32_t* seg_of, const uint32_t* seg_start, const uint32_t* seg_size,
                           uint32_t nseg, hipStream_t s) {
    if (!nseg) return hipSuccess;
    hipLaunchKernelGGL(k_seg_fill, dim3(nseg), dim3(256), 0, s, seg_of, seg_start, seg_size, nseg);
    return hipGetLastError();
}
hipError_t launch_seg_count_left(const uint8_t* side, const uint32_t* seg_of, uint64_t total,
                                 uint32_t* left, hipStream_t s) {
    uint64_t th = (total + 15) / 16;
    hipLaunchKernelGGL(k_seg_count_left, dim3((unsigned)((th + 255) / 256)), dim3(256), 0, s, side, seg_of, total,
                       left);
    return hipGetLastError();
}
hipError_t launch_seg_redeal(uint8_t* side, const uint32_t* seg_of, const uint8_t* degen, uint64_t total,
                             uint64_t seed, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_redeal, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, side, seg_of, degen,
                       total, seed);
    return hipGetLastError();
}
hipError_t launch_left_flags(const uint8_t* side, const uint32_t* seg_of, uint64_t total, uint32_t* flags,
                             hipStream_t s) {
    hipLaunchKernelGGL(k_left_flags, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, side, seg_of, total,
                       flags);
    return hipGetLastError();
}
hipError_t launch_seg_partition(const uint32_t* perm, uint32_t* perm_out, const uint8_t* side,
                                const uint32_t* seg_of, const uint32_t* seg_start, const uint32_t* left,
                                const uint32_t* lscan, uint64_t total, hipStream_t s) {
    hipLaunchKernelGGL(k_seg_partition, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, perm, perm_out, side,
                       seg_of, seg_start, left, lscan, total);
    return hipGetLastError();
}
