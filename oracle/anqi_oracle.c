/*
 * anqi_oracle.c -- CPU restatement of the ANQI approximate reverse-kNN path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  The product never links it.
 * Compiled with -ffp-contract=off: every fused multiply-add that the
 * reference's GCC build contracts (-ffp-contract=fast is GCC's default, the
 * reference Makefile does not change it) is spelled out with fmaf()/fma().
 *
 * Pinning (what checks each function; tests/test_oracle_*.py):
 *   or_ip/or_l2/or_norm/or_d2     float bits vs the reference's own functions
 *                                 compiled from include/distance.h (harness in
 *                                 oracle/ref_harness.cpp), random vectors,
 *                                 d in {1..300}; closed form on integer data.
 *   or_exact_topk                 dataset_gt golden (_baseknn_gt.bin) bit-exact;
 *                                 brute force sort on tiny n.
 *   or_rknn_membership            dataset_gt golden CSR for k=10/50/100.
 *   or_heap_push                  harness HeapList::push_nolock on random
 *                                 streams (same storage order).
 *   or_queue_*                    harness NeighborPriorityQueue streams.
 *   or_tree_entry, or_range_search  harness nndindex::range_search on the
 *                                 reference graph (per-query candidate lists,
 *                                 identical order) and #dist.
 *   or_lift / or_lift_query       reference _approxlift file + graph meta
 *                                 (bit-exact, single-thread reduction).
 *   or_recheck_key                harness recheck key bits.
 *   or_lrq_*                      reference ANQIVF2 artifact bytes (coeff and
 *                                 codes) and harness materialize_k thresholds.
 *   or_warm_start, or_nndescent_seq, or_gather, or_pool_topk, or_prune,
 *   or_commit                     harness calls into the reference's
 *                                 HeapList/nndindex methods on identical inputs
 *                                 (1 thread), bit-exact.
 *   inner-product metric ("ip"):
 *   or_ipdist                     float bits vs the reference's DistCal "ip"
 *                                 distance (harness), random vectors, d 1..300
 *                                 (the auto-vectorised loop of GCC 11.4 at the
 *                                 reference's flags; another compiler's
 *                                 vector-width choice moves the fused tail).
 *   or_lift_metric(ip)            the reference rknn_query's _approxlift file
 *                                 and graph meta for the "ip" sample
 *                                 (tests/golden/sample_ip), bit-exact.
 *   or_query_lift, or_recheck_key_ip, or_lrq_materialize_metric(ip)
 *                                 harness restatements of the reference's
 *                                 expressions (same flags); the reference
 *                                 rknn_query's stdout (recall, precision,
 *                                 candidates, #dist) on the "ip" sample.
 *   or_exact_topk_metric / or_rknn_membership_metric (ip)
 *                                 the reference dataset_gt's "ip" files for
 *                                 the sample, bit-exact.
 *   or_nndescent_snapshot         not a reference schedule (the reference
 *                                 LocalJoin is lock-order dependent); pinned
 *                                 by invariants: every applied push obeys the
 *                                 heap rule, heaps only improve, recall vs the
 *                                 sequential schedule within tolerance.
 */
#include "anqi_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_EPS 1e-6f   /* include/distance.h:14 */
static uint32_t prune_sel(const float* x, const float* nrm, uint32_t d, uint32_t base, uint32_t n, float* pd,
                          uint32_t* pi, uint32_t m, uint32_t M, float alpha, uint32_t* out);

/* GCC _mm512_reduce_add_ps tree (avx512fintrin.h, used by distance.h):
 * T3[i]=a[i+8]+a[i]; T6[i]=T3[i+4]+T3[i]; T8=[T6[0]+T6[2], T6[1]+T6[3]];
 * result T8[0]+T8[1]. */
static inline float tree16(const float* a) {
    float t3[8], t6[4], t80, t81;
    for (int i = 0; i < 8; i++) t3[i] = a[i + 8] + a[i];
    for (int i = 0; i < 4; i++) t6[i] = t3[i + 4] + t3[i];
    t80 = t6[0] + t6[2];
    t81 = t6[1] + t6[3];
    return t80 + t81;
}

/* distance.h:22-49 / 89-115: 16 FMA lanes over full chunks, tree reduce,
 * then the masked tail reduced (products, no FMA) and added. */
float or_ip(const float* x, const float* y, uint32_t d) {
    float acc[16];
    for (int l = 0; l < 16; l++) acc[l] = 0.0f;
    uint32_t m = d / 16, r = d % 16;
    for (uint32_t t = 0; t < m; t++)
        for (int l = 0; l < 16; l++) acc[l] = fmaf(x[16 * t + l], y[16 * t + l], acc[l]);
    float res = tree16(acc);
    if (r) {
        float p[16];
        for (int l = 0; l < 16; l++) p[l] = 0.0f;
        for (uint32_t l = 0; l < r; l++) p[l] = x[16 * m + l] * y[16 * m + l];
        res += tree16(p);
    }
    return res;
}

/* distance.h:58-86 */
float or_l2(const float* x, const float* y, uint32_t d) {
    float acc[16];
    for (int l = 0; l < 16; l++) acc[l] = 0.0f;
    uint32_t m = d / 16, r = d % 16;
    for (uint32_t t = 0; t < m; t++)
        for (int l = 0; l < 16; l++) {
            float df = x[16 * t + l] - y[16 * t + l];
            acc[l] = fmaf(df, df, acc[l]);
        }
    float res = tree16(acc);
    if (r) {
        float p[16];
        for (int l = 0; l < 16; l++) p[l] = 0.0f;
        for (uint32_t l = 0; l < r; l++) {
            float df = x[16 * m + l] - y[16 * m + l];
            p[l] = df * df;
        }
        res += tree16(p);
    }
    return res;
}

/* distance.h:143-161 (same structure as ip(x, x)). */
float or_norm(const float* x, uint32_t d) { return or_ip(x, x, d); }

/* index.h:172-176.  na + nb - 2*ip: 2*ip is exact, so the contracted
 * fnmadd and the separate subtract give the same bits. */
float or_d2(const float* x, const float* y, float nx, float ny, uint32_t d) {
    float ip = or_ip(x, y, d);
    return fmaf(-2.0f, ip, nx + ny);
}

void or_norms(const float* x, uint64_t n, uint32_t d, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) out[i] = or_norm(x + (size_t)i * d, d);
}

/* ---------------------------------------------------------------------- */
static int pair_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

static void sift(float* d, uint32_t* id, uint32_t m, uint32_t root) {
    for (;;) {
        uint32_t l = 2 * root + 1, big = root;
        if (l < m && pair_less(d[big], id[big], d[l], id[l])) big = l;
        if (l + 1 < m && pair_less(d[big], id[big], d[l + 1], id[l + 1])) big = l + 1;
        if (big == root) return;
        float td = d[root]; d[root] = d[big]; d[big] = td;
        uint32_t ti = id[root]; id[root] = id[big]; id[big] = ti;
        root = big;
    }
}

/* Heap sort by (dist, id) ascending; deterministic, no allocation. */
void or_sort_pairs(float* d, uint32_t* id, uint32_t m) {
    if (m < 2) return;
    for (int64_t r = (int64_t)m / 2 - 1; r >= 0; r--) sift(d, id, m, (uint32_t)r);
    for (uint32_t end = m - 1; end > 0; end--) {
        float td = d[0]; d[0] = d[end]; d[end] = td;
        uint32_t ti = id[0]; id[0] = id[end]; id[end] = ti;
        sift(d, id, end, 0);
    }
}

/* dataset_gt.cpp:206-245: key = fnmadd(2, ip, bn_i + bn_j); strict-< max-heap
 * over ascending j gives the K smallest (key, id) pairs. */
int or_exact_topk(const float* base, uint32_t n, uint32_t d, uint32_t K,
                  uint32_t* ids, float* keys, int threads) {
    return or_exact_topk_metric(base, n, d, K, 0, ids, keys, threads);
}

/* dataset_gt.cpp:206-245 with the metric: key = -ip for "ip" (:225). */
int or_exact_topk_metric(const float* base, uint32_t n, uint32_t d, uint32_t K, int ip_metric,
                         uint32_t* ids, float* keys, int threads) {
    if (K >= n) return -1;
    float* bn = (float*)malloc(sizeof(float) * n);
    or_norms(base, n, d, bn);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel
    {
        float* kd = (float*)malloc(sizeof(float) * (K + 1));
        uint32_t* ki = (uint32_t*)malloc(sizeof(uint32_t) * (K + 1));
#pragma omp for schedule(dynamic, 16)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            uint32_t i = (uint32_t)ii, cnt = 0;
            const float* xi = base + (size_t)i * d;
            /* bounded sorted insertion of (key, j) keeping the K smallest */
            for (uint32_t j = 0; j < n; j++) {
                if (j == i) continue;
                float ip = or_ip(xi, base + (size_t)j * d, d);
                float key = ip_metric ? -ip : fmaf(-2.0f, ip, bn[i] + bn[j]);
                if (cnt == K && !pair_less(key, j, kd[K - 1], ki[K - 1])) continue;
                uint32_t p = cnt < K ? cnt : K - 1;
                while (p > 0 && pair_less(key, j, kd[p - 1], ki[p - 1])) {
                    kd[p] = kd[p - 1]; ki[p] = ki[p - 1]; p--;
                }
                kd[p] = key; ki[p] = j;
                if (cnt < K) cnt++;
            }
            memcpy(ids + (size_t)i * K, ki, sizeof(uint32_t) * K);
            memcpy(keys + (size_t)i * K, kd, sizeof(float) * K);
        }
        free(kd); free(ki);
    }
    free(bn);
    return 0;
}

/* dataset_gt.cpp:283-313: hit iff qn + bn_j - 2 ip <= rk2[j]; rows sorted by
 * (key, id). */
uint64_t or_rknn_membership(const float* base, uint32_t n, const float* query,
                            uint32_t nq, uint32_t d, const float* rk2,
                            uint32_t* off, uint32_t* ids, float* keys, int threads) {
    return or_rknn_membership_metric(base, n, query, nq, d, 0, rk2, off, ids, keys, threads);
}

/* dataset_gt.cpp:283-313 with the metric: key = -ip for "ip" (:292). */
uint64_t or_rknn_membership_metric(const float* base, uint32_t n, const float* query,
                                   uint32_t nq, uint32_t d, int ip_metric, const float* rk2,
                                   uint32_t* off, uint32_t* ids, float* keys, int threads) {
    float* bn = (float*)malloc(sizeof(float) * n);
    or_norms(base, n, d, bn);
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    uint32_t* cnt = (uint32_t*)calloc(nq, sizeof(uint32_t));
#pragma omp parallel
    {
        float* hd = (float*)malloc(sizeof(float) * n);
        uint32_t* hi = (uint32_t*)malloc(sizeof(uint32_t) * n);
#pragma omp for schedule(dynamic, 4)
        for (int64_t qq = 0; qq < (int64_t)nq; qq++) {
            const float* q = query + (size_t)qq * d;
            float qn = or_norm(q, d);
            uint32_t c = 0;
            for (uint32_t j = 0; j < n; j++) {
                float ip = or_ip(q, base + (size_t)j * d, d);
                float key = ip_metric ? -ip : fmaf(-2.0f, ip, qn + bn[j]);
                if (key <= rk2[j]) { hd[c] = key; hi[c] = j; c++; }
            }
            cnt[qq] = c;
            if (ids) {
                or_sort_pairs(hd, hi, c);
                memcpy(ids + off[qq], hi, sizeof(uint32_t) * c);
                memcpy(keys + off[qq], hd, sizeof(float) * c);
            }
        }
        free(hd); free(hi);
    }
    uint64_t total = 0;
    if (!ids) {
        off[0] = 0;
        for (uint32_t q = 0; q < nq; q++) { total += cnt[q]; off[q + 1] = (uint32_t)total; }
    } else {
        total = off[nq];
    }
    free(cnt); free(bn);
    return total;
}

/* heap.h:83-186 (push_nolock == check_push with update_reverse=false). */
int or_heap_push(uint32_t* idx, float* val, uint32_t cap, uint32_t node, float value) {
    if (value > val[0]) return 0;
    for (uint32_t i = 0; i < cap; i++)
        if (idx[i] == node) return 0;
    uint32_t cur = 0, sw;
    for (;;) {
        uint32_t l = 2 * cur + 1, r = l + 1;
        if (l >= cap) break;
        else if (r >= cap) { if (val[l] > value) sw = l; else break; }
        else if (val[l] >= val[r]) { if (val[l] > value) sw = l; else break; }
        else { if (val[r] > value) sw = r; else break; }
        idx[cur] = idx[sw]; val[cur] = val[sw]; cur = sw;
    }
    idx[cur] = node; val[cur] = value;
    return 1;
}

/* neighbor.h:43-141 */
void or_queue_init(or_queue* q, float* dist, uint32_t* pid, uint32_t cap) {
    q->dist = dist; q->pid = pid; q->size = 0; q->cap = cap; q->cur = 0;
}

#define QID(p) ((p) & 0x7FFFFFFFu)

int or_queue_insert(or_queue* q, uint32_t id, float dist) {
    if (q->cap == 0) return 0;
    if (q->size == q->cap &&
        pair_less(q->dist[q->size - 1], QID(q->pid[q->size - 1]), dist, id)) return 0;
    uint32_t lo = 0, hi = q->size;
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (pair_less(dist, id, q->dist[mid], QID(q->pid[mid]))) hi = mid;
        else if (QID(q->pid[mid]) == id) return 0;
        else lo = mid + 1;
    }
    if (lo < q->cap) {
        memmove(q->dist + lo + 1, q->dist + lo, sizeof(float) * (q->size - lo));
        memmove(q->pid + lo + 1, q->pid + lo, sizeof(uint32_t) * (q->size - lo));
    }
    q->dist[lo] = dist; q->pid[lo] = id;
    if (q->size < q->cap) q->size++;
    if (lo < q->cur) q->cur = lo;
    return 1;
}

uint32_t or_queue_pop_closest(or_queue* q) {
    q->pid[q->cur] |= 0x80000000u;
    uint32_t pre = q->cur;
    while (q->cur < q->size && (q->pid[q->cur] >> 31)) q->cur++;
    return QID(q->pid[pre]);
}

/* rptrees.h:108-123.  Ties (|margin| < EPS) go left here; the reference
 * draws a random side, a measure-zero event on continuous data. */
uint32_t or_tree_entry(const or_tree* t, const float* q) {
    uint64_t k = t->nodes - 1;
    while (t->left[k] != (uint64_t)UINT32_MAX) {
        float margin = or_ip(t->hyper + (size_t)k * t->d, q, t->d) - t->offset[k];
        if (margin < -OR_EPS) k = t->left[k];
        else if (margin > OR_EPS) k = t->right[k];
        else k = t->left[k];
    }
    if (t->idx_off[k + 1] == t->idx_off[k]) return 0;
    return t->idx[t->idx_off[k]];
}

/* MT19937 (Matsumoto & Nishimura 1998): std::mt19937's seed() reduces the
 * seed mod 2^32 and runs the published initialisation recurrence. */
void or_mt_seed(or_mt* m, uint64_t seed) {
    m->x[0] = (uint32_t)seed;
    for (uint32_t i = 1; i < 624; i++) m->x[i] = 1812433253u * (m->x[i - 1] ^ (m->x[i - 1] >> 30)) + i;
    m->i = 624;
}

uint32_t or_mt_next(or_mt* m) {
    if (m->i >= 624) {
        for (uint32_t k = 0; k < 624; k++) {
            uint32_t y = (m->x[k] & 0x80000000u) | (m->x[(k + 1) % 624] & 0x7fffffffu);
            m->x[k] = m->x[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        m->i = 0;
    }
    uint32_t y = m->x[m->i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

/* Lemire, "Fast Random Integer Generation in an Interval" (TOMACS 2019): the
 * 64-bit product of a 32-bit draw and the range; reject while its low word is
 * below 2^32 mod range. */
uint32_t or_uniform_below(or_mt* m, uint32_t range) {
    uint64_t p = (uint64_t)or_mt_next(m) * range;
    uint32_t low = (uint32_t)p;
    if (low < range) {
        uint32_t t = (uint32_t)(0u - range) % range;
        while (low < t) {
            p = (uint64_t)or_mt_next(m) * range;
            low = (uint32_t)p;
        }
    }
    return (uint32_t)(p >> 32);
}

void or_mt_stream(uint64_t seed, uint32_t* out, uint64_t count) {
    or_mt m;
    or_mt_seed(&m, seed);
    for (uint64_t i = 0; i < count; i++) out[i] = or_mt_next(&m);
}

typedef struct {
    const float* data;
    uint32_t d, leaf, node_cap, nodes;
    or_mt mt;
    uint64_t draws, nidx;
    uint64_t *left, *right, *idx_off;
    float *offset, *hyper, *margin;
    uint32_t *idx, *scratch;
    int overflow;
} or_tb;

/* or_uniform_below, counting the engine outputs it consumes. */
static uint32_t tb_draw(or_tb* b, uint32_t range) {
    uint64_t p = (uint64_t)or_mt_next(&b->mt) * range;
    b->draws++;
    uint32_t low = (uint32_t)p;
    if (low < range) {
        uint32_t t = (uint32_t)(0u - range) % range;
        while (low < t) {
            p = (uint64_t)or_mt_next(&b->mt) * range;
            b->draws++;
            low = (uint32_t)p;
        }
    }
    return (uint32_t)(p >> 32);
}

static uint32_t tb_add(or_tb* b, uint64_t l, uint64_t r, float off, const float* hp,
                       const uint32_t* ids, uint64_t m) {
    if (b->nodes >= b->node_cap) { b->overflow = 1; return 0; }
    uint32_t k = b->nodes++;
    b->left[k] = l;
    b->right[k] = r;
    b->offset[k] = off;
    if (hp) memcpy(b->hyper + (size_t)k * b->d, hp, 4ull * b->d);
    else memset(b->hyper + (size_t)k * b->d, 0, 4ull * b->d);
    if (ids) memcpy(b->idx + b->nidx, ids, 4ull * m);
    b->nidx += m;
    b->idx_off[k + 1] = b->nidx;
    return k;
}

/* build_rptree_recursively (rptrees.h:127-152) with random_projection_split
 * (rptrees.h:28-85), L2 branch.  seg holds the node's indices in order; it is
 * partitioned stably in place (left part then right part). */
static void tb_rec(or_tb* b, uint32_t* seg, uint64_t m, uint32_t depth_left) {
    if (b->overflow) return;
    if (m <= b->leaf) { tb_add(b, UINT32_MAX, UINT32_MAX, FLT_MAX, NULL, seg, m); return; }
    if (depth_left == 0) {
        tb_add(b, UINT32_MAX, UINT32_MAX, FLT_MAX, NULL, seg, b->leaf < m ? b->leaf : m);
        return;
    }
    const uint32_t d = b->d;
    uint32_t r0 = tb_draw(b, (uint32_t)m), r1 = tb_draw(b, (uint32_t)m);
    if (r0 == r1) r0 = (uint32_t)((r1 + 1) % m);
    const float* a = b->data + (size_t)seg[r0] * d;
    const float* c = b->data + (size_t)seg[r1] * d;
    float* hp = (float*)malloc(4ull * d);
    float* mid = (float*)malloc(4ull * d);
    for (uint32_t j = 0; j < d; j++) {
        mid[j] = (a[j] + c[j]) / 2;
        hp[j] = a[j] - c[j];
    }
    float off = or_ip(hp, mid, d);
    free(mid);
    uint8_t* side = (uint8_t*)malloc(m);
    uint64_t nl = 0;
    for (uint64_t i = 0; i < m; i++) {
        float mg = or_ip(hp, b->data + (size_t)seg[i] * d, d) - off;
        if (mg < -OR_EPS) side[i] = 0;
        else if (mg > OR_EPS) side[i] = 1;
        else side[i] = tb_draw(b, 2) == 0 ? 0 : 1;
        nl += side[i] == 0;
    }
    if (nl == 0 || nl == m) {
        nl = 0;
        for (uint64_t i = 0; i < m; i++) {
            side[i] = tb_draw(b, 2) == 0 ? 0 : 1;
            nl += side[i] == 0;
        }
    }
    uint32_t* tmp = b->scratch;
    uint64_t pl = 0, pr = nl;
    for (uint64_t i = 0; i < m; i++) tmp[side[i] ? pr++ : pl++] = seg[i];
    memcpy(seg, tmp, 4ull * m);
    free(side);
    tb_rec(b, seg, nl, depth_left - 1);
    uint32_t lk = b->nodes - 1;
    tb_rec(b, seg + nl, m - nl, depth_left - 1);
    uint32_t rk = b->nodes - 1;
    if (!b->overflow) tb_add(b, lk, rk, off, hp, NULL, 0);
    free(hp);
}

uint32_t or_rp_tree_build(const float* data, uint32_t n, uint32_t d, uint32_t leaf,
                          uint32_t max_depth, uint64_t seed, uint32_t node_cap,
                          uint64_t* left, uint64_t* right, float* offset, float* hyper,
                          uint64_t* idx_off, uint32_t* idx, uint64_t* draws) {
    or_tb b;
    memset(&b, 0, sizeof(b));
    b.data = data; b.d = d; b.leaf = leaf; b.node_cap = node_cap;
    b.left = left; b.right = right; b.offset = offset; b.hyper = hyper;
    b.idx_off = idx_off; b.idx = idx;
    or_mt_seed(&b.mt, seed);
    uint32_t* seg = (uint32_t*)malloc(4ull * (n ? n : 1));
    b.scratch = (uint32_t*)malloc(4ull * (n ? n : 1));
    for (uint32_t i = 0; i < n; i++) seg[i] = i;
    idx_off[0] = 0;
    tb_rec(&b, seg, n, max_depth);
    free(seg);
    free(b.scratch);
    if (draws) *draws = b.draws;
    return b.overflow ? 0 : b.nodes;
}

/* index.h:1535-1652 with radj == NULL, eps = 0, CENTER score: the collect and
 * expand thresholds are both R2, compared as double(float D) <= R2. */
uint32_t or_range_search(const float* lifted, uint32_t n, uint32_t dl,
                         const uint32_t* deg, const uint32_t* links, uint32_t M,
                         uint32_t entry, const float* q, double R2, uint32_t L,
                         uint32_t* out, uint32_t cap, uint64_t* ndist) {
    uint8_t* vis = (uint8_t*)calloc(n, 1);
    float* qd = (float*)malloc(sizeof(float) * (L + 1));
    uint32_t* qp = (uint32_t*)malloc(sizeof(uint32_t) * (L + 1));
    uint32_t* flood = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t fsz = 0, fhead = 0, nres = 0;
    uint64_t nd = 0;
    or_queue beam;
    or_queue_init(&beam, qd, qp, L);
#define PUSH_RES(id) do { if (nres < cap) out[nres] = (id); nres++; } while (0)
    uint32_t ep = entry;
    float dep = or_l2(lifted + (size_t)ep * dl, q, dl); nd++;
    vis[ep] = 1;
    or_queue_insert(&beam, ep, dep);
    if ((double)dep <= R2) flood[fsz++] = ep;
    if ((double)dep <= R2) PUSH_RES(ep);
    while (beam.cur < beam.size) {
        uint32_t cur = or_queue_pop_closest(&beam);
        for (uint32_t j = 0; j < deg[cur]; j++) {
            uint32_t nb = links[(size_t)cur * M + j];
            if (vis[nb]) continue;
            vis[nb] = 1;
            float dd = or_l2(lifted + (size_t)nb * dl, q, dl); nd++;
            or_queue_insert(&beam, nb, dd);
            if ((double)dd <= R2) flood[fsz++] = nb;
            if ((double)dd <= R2) PUSH_RES(nb);
        }
    }
    while (fhead < fsz) {
        uint32_t cur = flood[fhead++];
        for (uint32_t j = 0; j < deg[cur]; j++) {
            uint32_t nb = links[(size_t)cur * M + j];
            if (vis[nb]) continue;
            vis[nb] = 1;
            float dd = or_l2(lifted + (size_t)nb * dl, q, dl); nd++;
            if ((double)dd <= R2) PUSH_RES(nb);
            if ((double)dd <= R2) flood[fsz++] = nb;
        }
    }
#undef PUSH_RES
    free(vis); free(qd); free(qp); free(flood);
    if (ndist) *ndist = nd;
    return nres;
}

/* rknn_query.cpp:2719-2730 (GRAPH_RANKM == 0, L2). */
void or_lift_query(const float* q, uint32_t d, double scale, float* qb) {
    for (uint32_t j = 0; j < d; j++) qb[j] = (float)((double)q[j] * scale);
    qb[d] = 1.0f;
    qb[d + 1] = 0.0f;
}

/* rknn_query.cpp:2719-2730 (GRAPH_RANKM == 0), both metrics: the lifted
 * query and the range-search bound R2 (v_j = double(q_j) * scale,
 * qn2s = fma(v_j, v_j, qn2s) in order; R2 = (qn2s + 1) + M*^2 for "ip",
 * M*^2 + 1 for "l2"). */
double or_query_lift(const float* q, uint32_t d, double scale, double mstar2, int ip_metric, float* qb) {
    double qn2s = 0.0;
    for (uint32_t j = 0; j < d; j++) {
        double v = (double)q[j] * scale;
        qb[j] = (float)v;
        qn2s = fma(v, v, qn2s);
    }
    qb[d] = 1.0f;
    qb[d + 1] = 0.0f;
    return ip_metric ? (qn2s + 1.0) + mstar2 : mstar2 + 1.0;
}

/* rknn_query.cpp:2555-2561, "ip": key = -<o, q>. */
float or_recheck_key_ip(const float* o, const float* q, uint32_t d) { return -or_ip(o, q, d); }

/* include/distance.h:170-175: nndindex's "ip" distance 1 - sum_i x_i y_i, a
 * scalar loop accumulated in element order from 0 in FP32.  GCC 11.4 at -O3
 * (-march=x86-64-v4, the reference Makefile's flags otherwise) vectorises the
 * products -- 16 at a time, then one block of 8 if at least 8 remain -- and
 * adds them to the sum one by one in element order (the reduction is kept in
 * order); the last d mod 8 elements run as scalar fused multiply-adds
 * (-ffp-contract=fast).  So: p += fl(x_i y_i) for i < nv, p = fma(x_i, y_i, p)
 * after, nv = 16 floor(d / 16) + (8 if d mod 16 >= 8). */
float or_ipdist(const float* x, const float* y, uint32_t d) {
    uint32_t nv = 16u * (d / 16u);
    if (d - nv >= 8u) nv += 8u;
    float p = 0.0f;
    for (uint32_t i = 0; i < nv; i++) {
        float t = x[i] * y[i];
        p = p + t;
    }
    for (uint32_t i = nv; i < d; i++) p = fmaf(x[i], y[i], p);
    return 1.0f - p;
}

/* rknn_query.cpp:2555-2561 */
float or_recheck_key(const float* o, const float* q, uint32_t d) {
    float ip = or_ip(o, q, d);
    float on = or_ip(o, o, d);
    float qn = or_ip(q, q, d);
    return fmaf(-2.0f, ip, on + qn);
}

/* rknn_query.cpp:482-513, policy "floor": upper_bound on float(value). */
uint32_t or_lrq_encode_value(double v, const float* centers, uint32_t code_size) {
    float fv = (float)v;
    uint32_t lo = 0, hi = code_size;   /* first index with centers[i] > fv */
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (fv < centers[mid]) hi = mid; else lo = mid + 1;
    }
    return lo == 0 ? 0 : lo - 1;
}

/* rknn_query.cpp:679-720 (bits 8).  GCC contracts the products into fma. */
void or_lrq_encode(const float* cutoffs, uint64_t n, uint32_t nk, uint32_t rankm,
                   const float* mu, const float* basis, const float* codebook,
                   uint32_t code_size, float* coeff, uint8_t* codes) {
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const float* rk = cutoffs + (size_t)ii * nk;
        float* ci = coeff + (size_t)ii * rankm;
        for (uint32_t m = 0; m < rankm; m++) {
            const float* gm = basis + (size_t)m * nk;
            double c = 0.0;
            for (uint32_t k = 0; k < nk; k++)
                c = fma((double)rk[k] - (double)mu[k], (double)gm[k], c);
            ci[m] = (float)c;
        }
        for (uint32_t k = 0; k < nk; k++) {
            double recon = (double)mu[k];
            for (uint32_t m = 0; m < rankm; m++)
                recon = fma((double)ci[m], (double)basis[(size_t)m * nk + k], recon);
            double residual = (double)rk[k] - recon;
            codes[(size_t)ii * nk + k] =
                (uint8_t)or_lrq_encode_value(residual, codebook + (size_t)k * code_size, code_size);
        }
    }
}

/* rknn_query.cpp:722-753 (bits 8, L2). */
void or_lrq_materialize(const float* coeff, const uint8_t* codes, uint64_t n,
                        uint32_t nk, uint32_t rankm, const float* mu,
                        const float* basis, const float* codebook,
                        uint32_t code_size, uint32_t k_one_based, double slack,
                        float* thr) {
    or_lrq_materialize_metric(coeff, codes, n, nk, rankm, mu, basis, codebook, code_size, k_one_based, slack, 0,
                              thr);
}

/* rknn_query.cpp:722-753 (bits 8); "ip" keeps the sign (:751). */
void or_lrq_materialize_metric(const float* coeff, const uint8_t* codes, uint64_t n,
                               uint32_t nk, uint32_t rankm, const float* mu,
                               const float* basis, const float* codebook,
                               uint32_t code_size, uint32_t k_one_based, double slack,
                               int ip_metric, float* thr) {
    uint32_t k = k_one_based - 1;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        const float* ci = coeff + (size_t)ii * rankm;
        double value = (double)mu[k];
        for (uint32_t m = 0; m < rankm; m++)
            value = fma((double)ci[m], (double)basis[(size_t)m * nk + k], value);
        value += (double)codebook[(size_t)k * code_size + codes[(size_t)ii * nk + k]];
        double scaled = slack * value;
        thr[ii] = ip_metric ? (float)scaled : (float)(0.0 < scaled ? scaled : 0.0);
    }
}

/* rknn_query.cpp:2203-2236 (GRAPH_RANKM == 0, L2, postfilter horizon). */
void or_lift(const float* base, uint64_t n, uint32_t d, const float* rk_horizon,
             float* lifted, double* mstar2, double* scale) {
    or_lift_metric(base, n, d, rk_horizon, 0, lifted, mstar2, scale);
}

/* rknn_query.cpp:2203-2236 (GRAPH_RANKM == 0, postfilter horizon), both
 * metrics: the lift coordinate is float(r s^2) for "ip" (:2227). */
void or_lift_metric(const float* base, uint64_t n, uint32_t d, const float* rk_horizon, int ip_metric,
                    float* lifted, double* mstar2, double* scale) {
    double sum_n2 = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        const float* o = base + i * d;
        double s = 0.0;
        for (uint32_t j = 0; j < d; j++) s = fma((double)o[j], (double)o[j], s);
        sum_n2 += s;
    }
    double s_scale = 1.0 / sqrt(sum_n2 / (double)n + 1e-12), sc2 = s_scale * s_scale;
    uint32_t dl = d + 1, dout = d + 2;
    double maxn2 = 0.0;
    for (uint64_t i = 0; i < n; i++) {
        const float* o = base + i * d;
        float* h = lifted + i * dout;
        double nr2 = 0.0;
        for (uint32_t j = 0; j < d; j++) {
            double v = (double)o[j] * s_scale;
            h[j] = (float)v;
            nr2 = fma(v, v, nr2);
        }
        h[d] = ip_metric ? (float)((double)rk_horizon[i] * sc2)
                         : (float)(0.5 * fma((double)rk_horizon[i], sc2, -nr2));
        double hn2 = fma((double)h[d], (double)h[d], nr2);
        if (hn2 > maxn2) maxn2 = hn2;
    }
    double M2 = maxn2 * 1.0001;
    for (uint64_t i = 0; i < n; i++) {
        float* h = lifted + i * dout;
        double hn2 = 0.0;
        for (uint32_t j = 0; j < dl; j++) hn2 = fma((double)h[j], (double)h[j], hn2);
        double pad = M2 - hn2;
        if (pad < 0) pad = 0;
        h[dl] = (float)sqrt(pad);
    }
    *mstar2 = M2;
    *scale = s_scale;
}

/* index.h:1204-1218 (flat warm start). */
void or_warm_start(const float* x, const float* nrm, uint32_t n, uint32_t d,
                   const uint32_t* ids, uint32_t width, uint32_t cap,
                   uint32_t* hidx, float* hval) {
    for (size_t t = 0; t < (size_t)n * cap; t++) { hidx[t] = UINT32_MAX; hval[t] = FLT_MAX; }
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        uint32_t i = (uint32_t)ii;
        for (uint32_t k = 0; k < width; k++) {
            uint32_t j = ids[(size_t)i * width + k];
            if (j != i && j < n)
                or_heap_push(hidx + (size_t)i * cap, hval + (size_t)i * cap, cap, j,
                             or_d2(x + (size_t)i * d, x + (size_t)j * d, nrm[i], nrm[j], d));
        }
    }
}

static uint32_t take_sample(const uint32_t* row, uint32_t cap, int sample, uint32_t* loc) {
    uint32_t c = 0;
    for (uint32_t s = 0; s < cap; s++) {
        if (row[s] != UINT32_MAX) { loc[c++] = row[s]; if ((int)c >= sample) break; }
    }
    return c;
}

/* index.h:1468-1487 with one thread. */
int or_nndescent_seq(const float* x, const float* nrm, uint32_t n, uint32_t d,
                     uint32_t cap, uint32_t* hidx, float* hval, int iters,
                     int sample, int64_t* upd) {
    uint32_t loc[128];
    if (sample > 120) sample = 120;
    int it;
    for (it = 0; it < iters; it++) {
        int64_t u_cnt = 0;
        for (uint32_t u = 0; u < n; u++) {
            uint32_t cnt = take_sample(hidx + (size_t)u * cap, cap, sample, loc);
            for (uint32_t a = 0; a < cnt; a++)
                for (uint32_t b = a + 1; b < cnt; b++) {
                    uint32_t v = loc[a], w = loc[b];
                    float dd = or_d2(x + (size_t)v * d, x + (size_t)w * d, nrm[v], nrm[w], d);
                    u_cnt += or_heap_push(hidx + (size_t)v * cap, hval + (size_t)v * cap, cap, w, dd);
                    u_cnt += or_heap_push(hidx + (size_t)w * cap, hval + (size_t)w * cap, cap, v, dd);
                }
        }
        if (upd) upd[it] = u_cnt;
        if (u_cnt < (int64_t)(n / 200)) { it++; break; }
    }
    return it;
}

typedef struct { uint32_t row, id; float d; } or_prop;

static int prop_cmp(const void* a, const void* b) {
    const or_prop* x = (const or_prop*)a;
    const or_prop* y = (const or_prop*)b;
    if (x->row != y->row) return x->row < y->row ? -1 : 1;
    if (x->d != y->d) return x->d < y->d ? -1 : 1;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    return 0;
}

int or_nndescent_snapshot(const float* x, const float* nrm, uint32_t n, uint32_t d,
                          uint32_t cap, uint32_t* hidx, float* hval, int iters,
                          int sample, int64_t* upd) {
    if (sample > 120) sample = 120;
    uint32_t* samp = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n * sample);
    uint32_t* scnt = (uint32_t*)malloc(sizeof(uint32_t) * n);
    int it;
    for (it = 0; it < iters; it++) {
        size_t np = 0;
        for (uint32_t u = 0; u < n; u++) {
            scnt[u] = take_sample(hidx + (size_t)u * cap, cap, sample, samp + (size_t)u * sample);
            np += (size_t)scnt[u] * (scnt[u] - 1);
        }
        or_prop* P = (or_prop*)malloc(sizeof(or_prop) * (np ? np : 1));
        size_t k = 0;
        for (uint32_t u = 0; u < n; u++) {
            const uint32_t* loc = samp + (size_t)u * sample;
            for (uint32_t a = 0; a < scnt[u]; a++)
                for (uint32_t b = a + 1; b < scnt[u]; b++) {
                    uint32_t v = loc[a], w = loc[b];
                    float dd = or_d2(x + (size_t)v * d, x + (size_t)w * d, nrm[v], nrm[w], d);
                    P[k].row = v; P[k].id = w; P[k].d = dd; k++;
                    P[k].row = w; P[k].id = v; P[k].d = dd; k++;
                }
        }
        qsort(P, np, sizeof(or_prop), prop_cmp);
        int64_t u_cnt = 0;
        for (size_t t = 0; t < np; t++)
            u_cnt += or_heap_push(hidx + (size_t)P[t].row * cap, hval + (size_t)P[t].row * cap,
                                  cap, P[t].id, P[t].d);
        free(P);
        if (upd) upd[it] = u_cnt;
        if (u_cnt < (int64_t)(n / 200)) { it++; break; }
    }
    free(samp); free(scnt);
    return it;
}

/* heap.h:190-207 in row order (the single-thread order). */
void or_reverse_csr(const uint32_t* hidx, uint32_t n, uint32_t cap,
                    uint32_t* roff, uint32_t* rids) {
    memset(roff, 0, sizeof(uint32_t) * (n + 1));
    for (size_t t = 0; t < (size_t)n * cap; t++)
        if (hidx[t] < n) roff[hidx[t] + 1]++;
    for (uint32_t i = 0; i < n; i++) roff[i + 1] += roff[i];
    uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * n);
    memcpy(cur, roff, sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t s = 0; s < cap; s++) {
            uint32_t j = hidx[(size_t)i * cap + s];
            if (j < n) rids[cur[j]++] = i;
        }
    free(cur);
}

/* index.h:1491-1525 */
uint32_t or_gather(const float* x, const float* nrm, uint32_t n, uint32_t d,
                   const uint32_t* hidx, const float* hval, uint32_t cap,
                   const uint32_t* roff, const uint32_t* rids, uint32_t i,
                   uint32_t explore, uint32_t pool_cap, float* pool_d,
                   uint32_t* pool_id, uint32_t pool_max) {
    uint8_t* vis = (uint8_t*)calloc(n, 1);
    float* qd = (float*)malloc(sizeof(float) * (explore + 1));
    uint32_t* qp = (uint32_t*)malloc(sizeof(uint32_t) * (explore + 1));
    or_queue cq;
    or_queue_init(&cq, qd, qp, explore);
    uint32_t m = 0;
#define PUSHP(id_, d_) do { if (m < pool_max) { pool_d[m] = (d_); pool_id[m] = (id_); } m++; } while (0)
    const float* xi = x + (size_t)i * d;
    for (uint32_t s = 0; s < cap; s++) {
        uint32_t id = hidx[(size_t)i * cap + s];
        if (id == UINT32_MAX || id >= n || vis[id]) continue;
        vis[id] = 1;
        float dv = hval[(size_t)i * cap + s];
        or_queue_insert(&cq, id, dv);
        PUSHP(id, dv);
    }
    while (cq.cur < cq.size) {
        uint32_t cur = or_queue_pop_closest(&cq);
        for (uint32_t s = 0; s < cap; s++) {
            uint32_t nb = hidx[(size_t)cur * cap + s];
            if (nb == UINT32_MAX || nb >= n || vis[nb]) continue;
            vis[nb] = 1;
            float dv = or_d2(xi, x + (size_t)nb * d, nrm[i], nrm[nb], d);
            or_queue_insert(&cq, nb, dv);
            PUSHP(nb, dv);
        }
        for (uint32_t p = roff[cur]; p < roff[cur + 1]; p++) {
            uint32_t nb = rids[p];
            if (nb == UINT32_MAX || nb >= n || vis[nb]) continue;
            vis[nb] = 1;
            float dv = or_d2(xi, x + (size_t)nb * d, nrm[i], nrm[nb], d);
            or_queue_insert(&cq, nb, dv);
            PUSHP(nb, dv);
        }
        if (m >= pool_cap) break;
    }
#undef PUSHP
    free(vis); free(qd); free(qp);
    return m;
}

uint32_t or_pool_topk(uint32_t self, const float* pool_d, const uint32_t* pool_id,
                      uint32_t m, uint32_t k, uint32_t* ids, float* keys) {
    float* d = (float*)malloc(sizeof(float) * (m ? m : 1));
    uint32_t* id = (uint32_t*)malloc(sizeof(uint32_t) * (m ? m : 1));
    memcpy(d, pool_d, sizeof(float) * m);
    memcpy(id, pool_id, sizeof(uint32_t) * m);
    or_sort_pairs(d, id, m);
    uint32_t c = 0;
    for (uint32_t t = 0; t < m && c < k; t++) {
        if (id[t] == self) continue;
        ids[c] = id[t]; keys[c] = d[t]; c++;
    }
    free(d); free(id);
    return c;
}

/* index.h:225-248 + 267-281 + default branch of 772-829. */
uint32_t or_prune(const float* x, const float* nrm, uint32_t d, uint32_t base,
                  uint32_t n, float* pool_d, uint32_t* pool_id, uint32_t m,
                  uint32_t M, float alpha, uint32_t* out) {
    /* drop sentinel/out-of-range/self, keep min dist per id */
    uint32_t w = 0;
    for (uint32_t t = 0; t < m; t++) {
        uint32_t id = pool_id[t];
        if (id == UINT32_MAX || id >= n || id == base) continue;
        pool_id[w] = id; pool_d[w] = pool_d[t]; w++;
    }
    /* sort by (id, dist), keep the minimum distance per id */
    {
        or_prop* P = (or_prop*)malloc(sizeof(or_prop) * (w ? w : 1));
        for (uint32_t t = 0; t < w; t++) { P[t].row = pool_id[t]; P[t].id = 0; P[t].d = pool_d[t]; }
        qsort(P, w, sizeof(or_prop), prop_cmp);   /* (row=id, d) order */
        uint32_t u = 0;
        for (uint32_t r = 0; r < w;) {
            uint32_t e = r + 1;
            while (e < w && P[e].row == P[r].row) e++;
            pool_id[u] = P[r].row; pool_d[u] = P[r].d; u++;   /* first = min d */
            r = e;
        }
        free(P);
        w = u;
    }
    or_sort_pairs(pool_d, pool_id, w);
    uint32_t deg = 0;
    for (int pass = 0; pass < 2; pass++) {   /* rng_extend(pool, M) twice */
        for (uint32_t t = 0; t < w; t++) {
            if (deg >= M) break;
            uint32_t cid = pool_id[t];
            float d2c = pool_d[t];
            int sel = 0;
            for (uint32_t s = 0; s < deg; s++) if (out[s] == cid) { sel = 1; break; }
            if (sel) continue;
            int occ = 0;
            for (uint32_t s = 0; s < deg; s++) {
                uint32_t sid = out[s];
                float dsc = or_d2(x + (size_t)sid * d, x + (size_t)cid * d, nrm[sid], nrm[cid], d);
                if (alpha * dsc < d2c) { occ = 1; break; }
            }
            if (!occ) out[deg++] = cid;
        }
    }
    return deg;
}

/* index.h:907-930 */
void or_commit(const float* x, const float* nrm, uint32_t n, uint32_t d,
               const uint32_t* O, const uint32_t* ocnt, uint32_t M, float alpha,
               uint32_t* links, uint32_t* deg) {
    uint32_t* roff = (uint32_t*)calloc(n + 1, sizeof(uint32_t));
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t t = 0; t < ocnt[i]; t++) roff[O[(size_t)i * M + t] + 1]++;
    for (uint32_t i = 0; i < n; i++) roff[i + 1] += roff[i];
    uint32_t* rids = (uint32_t*)malloc(sizeof(uint32_t) * (roff[n] ? roff[n] : 1));
    uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * n);
    memcpy(cur, roff, sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t t = 0; t < ocnt[i]; t++) { uint32_t j = O[(size_t)i * M + t]; rids[cur[j]++] = i; }
#pragma omp parallel
    {
        float* pd = (float*)malloc(sizeof(float) * (M + n));
        uint32_t* pi = (uint32_t*)malloc(sizeof(uint32_t) * (M + n));
#pragma omp for schedule(dynamic, 64)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            uint32_t i = (uint32_t)ii, m = 0;
            for (uint32_t t = 0; t < ocnt[i]; t++) {
                uint32_t j = O[(size_t)i * M + t];
                pd[m] = or_d2(x + (size_t)i * d, x + (size_t)j * d, nrm[i], nrm[j], d); pi[m] = j; m++;
            }
            for (uint32_t p = roff[i]; p < roff[i + 1]; p++) {
                uint32_t j = rids[p];
                pd[m] = or_d2(x + (size_t)i * d, x + (size_t)j * d, nrm[i], nrm[j], d); pi[m] = j; m++;
            }
            deg[i] = prune_sel(x, nrm, d, i, n, pd, pi, m, M, alpha, links + (size_t)i * M);
        }
        free(pd); free(pi);
    }
    free(roff); free(rids); free(cur);
}

/* ---- the prune used by commit_edges and the Vamana passes ---------------------- */
/* or_set_edge_policy: the edge policy every later or_commit / or_vamana_*
 * prune follows (NULL: the plain prune), as the reference's index applies its
 * configured policy to each prune_pool_to_ids. */
static const or_edge_policy* g_edge_policy = 0;
void or_set_edge_policy(const or_edge_policy* E) { g_edge_policy = E; }
static uint32_t prune_sel(const float* x, const float* nrm, uint32_t d, uint32_t base, uint32_t n, float* pd,
                          uint32_t* pi, uint32_t m, uint32_t M, float alpha, uint32_t* out) {
    if (g_edge_policy) return or_prune_policy(x, nrm, d, base, n, pd, pi, m, M, alpha, g_edge_policy, out);
    return or_prune(x, nrm, d, base, n, pd, pi, m, M, alpha, out);
}

/* ---- Vamana refinement and the single graph entry -------------------------- */

/* greedy_collect (index.h:831-855) and greedy_collect_from_entry (:857-880):
 * a beam of Lb entries from `entry` over the fixed-degree graph; every first
 * visit except row i itself joins the pool, in visit order. */
uint32_t or_vamana_collect(const float* x, const float* nrm, uint32_t n, uint32_t d,
                           const uint32_t* deg, const uint32_t* links, uint32_t M,
                           uint32_t i, uint32_t entry, uint32_t Lb,
                           float* pool_d, uint32_t* pool_id, uint32_t pool_max) {
    uint8_t* vis = (uint8_t*)calloc(n, 1);
    float* qd = (float*)malloc(sizeof(float) * (Lb + 1));
    uint32_t* qp = (uint32_t*)malloc(sizeof(uint32_t) * (Lb + 1));
    or_queue q;
    or_queue_init(&q, qd, qp, Lb);
    uint32_t m = 0;
    const float* xi = x + (size_t)i * d;
#define PUSHV(id_, d_) do { if ((id_) != i) { if (m < pool_max) { pool_d[m] = (d_); pool_id[m] = (id_); } m++; } } while (0)
    const uint32_t ep = entry < n ? entry : 0;
    vis[ep] = 1;
    {
        const float dv = or_d2(xi, x + (size_t)ep * d, nrm[i], nrm[ep], d);
        or_queue_insert(&q, ep, dv);
        PUSHV(ep, dv);
    }
    while (q.cur < q.size) {
        const uint32_t cur = or_queue_pop_closest(&q);
        for (uint32_t j = 0; j < deg[cur]; j++) {
            const uint32_t nb = links[(size_t)cur * M + j];
            if (nb >= n || vis[nb]) continue;
            vis[nb] = 1;
            const float dv = or_d2(xi, x + (size_t)nb * d, nrm[i], nrm[nb], d);
            or_queue_insert(&q, nb, dv);
            PUSHV(nb, dv);
        }
    }
#undef PUSHV
    free(vis); free(qd); free(qp);
    return m;
}

/* append_aknn_edges (index.h:766-769): the heap slots with their distances
 * recomputed; append_current_edges (:757-764). */
static uint32_t vam_append_heap(const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* hidx,
                                uint32_t cap, uint32_t i, float* pd, uint32_t* pi, uint32_t m) {
    for (uint32_t s = 0; s < cap; s++) {
        const uint32_t j = hidx[(size_t)i * cap + s];
        if (j == UINT32_MAX || j >= n || j == i) continue;
        pd[m] = or_d2(x + (size_t)i * d, x + (size_t)j * d, nrm[i], nrm[j], d);
        pi[m++] = j;
    }
    return m;
}
static uint32_t vam_append_links(const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* deg,
                                 const uint32_t* links, uint32_t M, uint32_t i, float* pd, uint32_t* pi,
                                 uint32_t m) {
    for (uint32_t j = 0; j < deg[i]; j++) {
        const uint32_t nb = links[(size_t)i * M + j];
        if (nb == UINT32_MAX || nb >= n || nb == i) continue;
        pd[m] = or_d2(x + (size_t)i * d, x + (size_t)nb * d, nrm[i], nrm[nb], d);
        pi[m++] = nb;
    }
    return m;
}

/* One pass of the global style (index.h:1267-1279): per row, collect from the
 * row's tree entry, add the heap slots, prune to M; then commit_edges. */
void or_vamana_global_pass(const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* entries,
                           const uint32_t* hidx, uint32_t cap, uint32_t Lb, uint32_t M, float alpha,
                           uint32_t* links, uint32_t* deg) {
    uint32_t* O = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n * M);
    uint32_t* ocnt = (uint32_t*)malloc(sizeof(uint32_t) * n);
#pragma omp parallel
    {
        float* pd = (float*)malloc(sizeof(float) * ((size_t)n + cap));
        uint32_t* pi = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + cap));
#pragma omp for schedule(dynamic, 16)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            const uint32_t i = (uint32_t)ii;
            uint32_t m = or_vamana_collect(x, nrm, n, d, deg, links, M, i, entries[i], Lb, pd, pi, n);
            m = vam_append_heap(x, nrm, n, d, hidx, cap, i, pd, pi, m);
            ocnt[i] = prune_sel(x, nrm, d, i, n, pd, pi, m, M, alpha, O + (size_t)i * M);
        }
        free(pd); free(pi);
    }
    or_commit(x, nrm, n, d, O, ocnt, M, alpha, links, deg);
    free(O); free(ocnt);
}

/* MT19937-64 (Nishimura 2000; std::mt19937_64's parameters). */
void or_mt64_seed(or_mt64* m, uint64_t seed) {
    m->x[0] = seed;
    for (uint32_t i = 1; i < 312; i++) m->x[i] = 6364136223846793005ull * (m->x[i - 1] ^ (m->x[i - 1] >> 62)) + i;
    m->i = 312;
}

uint64_t or_mt64_next(or_mt64* m) {
    if (m->i >= 312) {
        for (uint32_t k = 0; k < 312; k++) {
            const uint64_t y = (m->x[k] & 0xFFFFFFFF80000000ull) | (m->x[(k + 1) % 312] & 0x7FFFFFFFull);
            m->x[k] = m->x[(k + 156) % 312] ^ (y >> 1) ^ ((y & 1u) ? 0xB5026F5AA96619E9ull : 0ull);
        }
        m->i = 0;
    }
    uint64_t y = m->x[m->i++];
    y ^= (y >> 29) & 0x5555555555555555ull;
    y ^= (y << 17) & 0x71D67FFFEDA60000ull;
    y ^= (y << 37) & 0xFFF7EEE000000000ull;
    y ^= y >> 43;
    return y;
}

/* Lemire's nearly-divisionless draw in [0, range) from a 64-bit engine: the
 * 128-bit product's high word, rejecting while its low word is below
 * 2^64 mod range (libstdc++'s uniform_int_distribution for a full 64-bit
 * engine). */
static uint64_t mt64_below(or_mt64* m, uint64_t range) {
    unsigned __int128 p = (unsigned __int128)or_mt64_next(m) * range;
    uint64_t low = (uint64_t)p;
    if (low < range) {
        const uint64_t t = (0ull - range) % range;
        while (low < t) {
            p = (unsigned __int128)or_mt64_next(m) * range;
            low = (uint64_t)p;
        }
    }
    return (uint64_t)(p >> 64);
}

/* std::shuffle as libstdc++ (GCC 11 bits/stl_algo.h) runs it with a 64-bit
 * engine over n < 2^32 elements: one draw in [0, b(b+1)) gives the swap
 * positions of two successive elements (x / (b+1), x % (b+1)), after one
 * single draw in [0, 2) when n is even. */
void or_shuffle_u32(uint32_t* a, uint32_t n, or_mt64* g) {
    if (n == 0) return;
    uint32_t i = 1;
    uint32_t t;
    if ((n % 2u) == 0) {
        const uint64_t p = mt64_below(g, 2);
        t = a[i]; a[i] = a[p]; a[p] = t;
        i++;
    }
    while (i != n) {
        const uint64_t b = (uint64_t)i + 1;
        const uint64_t xr = mt64_below(g, b * (b + 1));
        const uint64_t p1 = xr / (b + 1), p2 = xr % (b + 1);
        t = a[i]; a[i] = a[p1]; a[p1] = t;
        i++;
        t = a[i]; a[i] = a[p2]; a[p2] = t;
        i++;
    }
}

/* choose_vamana_entry's "medoid" (index.h:190-222): the sampled row nearest
 * (in binary64, fused squares) to the sample's binary64 mean. */
uint32_t or_vamana_medoid(const float* x, uint32_t n, uint32_t d, uint64_t sample) {
    if (sample < 1) sample = 1;
    if (sample > n) sample = n;
    uint64_t step = n / sample;
    if (step < 1) step = 1;
    double* mean = (double*)calloc(d, sizeof(double));
    for (uint64_t t = 0; t < sample; t++) {
        const float* r = x + ((t * step) % n) * (size_t)d;
        for (uint32_t j = 0; j < d; j++) mean[j] += (double)r[j];
    }
    for (uint32_t j = 0; j < d; j++) mean[j] /= (double)sample;
    uint32_t best = 0;
    double bestd = INFINITY;
    for (uint64_t t = 0; t < sample; t++) {
        const uint64_t id = (t * step) % n;
        const float* r = x + id * (size_t)d;
        double dd = 0.0;
        for (uint32_t j = 0; j < d; j++) {
            const double z = (double)r[j] - mean[j];
            dd = fma(z, z, dd);
        }
        if (dd < bestd) { bestd = dd; best = (uint32_t)id; }
    }
    free(mean);
    return best;
}

/* vamana_batch_refine (index.h:932-1063): per pass (prune alpha early_alpha
 * before the last pass when use_early), the order reshuffled by
 * mt19937_64(seed + 0x9e3779b97f4a7c15 (pass + 1)); batches of 1, 2, 4, ...
 * rows (base^k - 1 .. base^(k+1) - 1 while base^k <= max_batch), then
 * max_batch rows.  A batch's rows collect from the one entry, add their heap
 * slots and current edges and are pruned against the graph as it stood
 * before the batch, then written; every row they point to is pruned again
 * over its new edges plus the batch rows that point to it. */
void or_vamana_batch(const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* hidx, uint32_t cap,
                     uint32_t Lb, int passes, double batch_frac, uint64_t max_batch_cap, double batch_base,
                     float alpha, float early_alpha, int use_early, uint32_t entry, uint64_t seed, uint32_t M,
                     uint32_t* links, uint32_t* deg) {
    if (n == 0 || passes <= 0) return;
    if (batch_frac <= 0.0) batch_frac = 0.02;
    double mb = batch_frac * (double)n;
    if (mb < 1.0) mb = 1.0;
    uint64_t max_batch = (uint64_t)mb;
    if (max_batch > max_batch_cap) max_batch = max_batch_cap;
    if (max_batch == 0) max_batch = n;
    if (batch_base <= 1.0) batch_base = 2.0;
    uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; i++) order[i] = i;
    uint32_t* O = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n * M);
    uint32_t* ocnt = (uint32_t*)malloc(sizeof(uint32_t) * n);
    uint32_t* rcnt = (uint32_t*)calloc(n, sizeof(uint32_t));
    uint32_t* roff = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    uint32_t* rsrc = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n * M);
    uint32_t* tl = (uint32_t*)malloc(sizeof(uint32_t) * n);
    for (int pass = 0; pass < passes; pass++) {
        const float pa = (use_early && pass + 1 < passes) ? early_alpha : alpha;
        or_mt64 g;
        or_mt64_seed(&g, seed + 0x9e3779b97f4a7c15ull * (uint64_t)(pass + 1));
        or_shuffle_u32(order, n, &g);
        uint64_t count = 0, inc = 0;
        while (count < n) {
            uint64_t fl = count, ce;
            const double p0 = pow(batch_base, (double)inc);
            if (p0 <= (double)max_batch) {
                fl = (uint64_t)p0 - 1;
                const double p1 = pow(batch_base, (double)(inc + 1));
                ce = (uint64_t)p1 - 1;
                if (ce > n) ce = n;
                count = ce;
            } else {
                ce = count + max_batch;
                if (ce > n) ce = n;
                count = ce;
            }
            inc++;
            if (ce <= fl || fl >= n) continue;
            const uint32_t bsz = (uint32_t)(ce - fl);
#pragma omp parallel
            {
                float* pd = (float*)malloc(sizeof(float) * ((size_t)n + cap + M));
                uint32_t* pi = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + cap + M));
#pragma omp for schedule(dynamic, 16)
                for (int64_t b = 0; b < (int64_t)bsz; b++) {
                    const uint32_t node = order[fl + b];
                    uint32_t m = or_vamana_collect(x, nrm, n, d, deg, links, M, node, entry, Lb, pd, pi, n);
                    m = vam_append_heap(x, nrm, n, d, hidx, cap, node, pd, pi, m);
                    m = vam_append_links(x, nrm, n, d, deg, links, M, node, pd, pi, m);
                    ocnt[node] = prune_sel(x, nrm, d, node, n, pd, pi, m, M, pa, O + (size_t)node * M);
                }
                free(pd); free(pi);
            }
            /* write_ids, then the rows they point to */
            uint32_t nt = 0;
            for (uint32_t b = 0; b < bsz; b++) {
                const uint32_t node = order[fl + b];
                for (uint32_t t = 0; t < ocnt[node]; t++) {
                    const uint32_t dst = O[(size_t)node * M + t];
                    links[(size_t)node * M + t] = dst;
                    if (dst < n && rcnt[dst]++ == 0) tl[nt++] = dst;
                }
                for (uint32_t t = ocnt[node]; t < M; t++) links[(size_t)node * M + t] = UINT32_MAX;
                deg[node] = ocnt[node];
            }
            roff[0] = 0;
            for (uint32_t t = 0; t < nt; t++) roff[t + 1] = roff[t] + rcnt[tl[t]];
            for (uint32_t t = 0; t < nt; t++) rcnt[tl[t]] = roff[t];   /* cursors */
            for (uint32_t b = 0; b < bsz; b++) {
                const uint32_t node = order[fl + b];
                for (uint32_t t = 0; t < ocnt[node]; t++) {
                    const uint32_t dst = O[(size_t)node * M + t];
                    if (dst < n) rsrc[rcnt[dst]++] = node;
                }
            }
#pragma omp parallel
            {
                float* pd = (float*)malloc(sizeof(float) * ((size_t)n + M));
                uint32_t* pi = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + M));
                uint32_t* out = (uint32_t*)malloc(sizeof(uint32_t) * M);
#pragma omp for schedule(dynamic, 16)
                for (int64_t t = 0; t < (int64_t)nt; t++) {
                    const uint32_t dst = tl[t];
                    uint32_t m = vam_append_links(x, nrm, n, d, deg, links, M, dst, pd, pi, 0);
                    for (uint32_t p = roff[t]; p < roff[t + 1]; p++) {
                        const uint32_t src = rsrc[p];
                        if (src == dst || src >= n) continue;
                        pd[m] = or_d2(x + (size_t)dst * d, x + (size_t)src * d, nrm[dst], nrm[src], d);
                        pi[m++] = src;
                    }
                    const uint32_t dg = prune_sel(x, nrm, d, dst, n, pd, pi, m, M, pa, out);
                    /* rows are pruned from copies, so the writes may follow in any order */
#pragma omp critical
                    {
                        for (uint32_t s = 0; s < M; s++) links[(size_t)dst * M + s] = s < dg ? out[s] : UINT32_MAX;
                        deg[dst] = dg;
                    }
                }
                free(pd); free(pi); free(out);
            }
            for (uint32_t t = 0; t < nt; t++) rcnt[tl[t]] = 0;
        }
    }
    free(order); free(O); free(ocnt); free(rcnt); free(roff); free(rsrc); free(tl);
}

/* ---- edge policies of prune_pool_to_ids (include/index.h:283-829) ---------- */

static float pol_d2(const float* x, const float* nrm, uint32_t d, uint32_t a, uint32_t b) {
    return or_d2(x + (size_t)a * d, x + (size_t)b * d, nrm[a], nrm[b], d);
}
static int pol_has(const uint32_t* ids, uint32_t m, uint32_t id) {
    for (uint32_t t = 0; t < m; t++) if (ids[t] == id) return 1;
    return 0;
}
/* edge_radj (:291-293): scale^2 times the residual of rank k */
static double pol_radj(const or_edge_policy* E, uint32_t id, int k) {
    return E->scale2 * (double)E->resid[(size_t)id * E->nk + (size_t)(k - 1)];
}
/* edge_sort_radj (:295-303) */
static double pol_sort_radj(const or_edge_policy* E, uint32_t id) {
    if (!E->ball || E->nks == 0) return 0.0;
    double best = -INFINITY;
    for (uint32_t t = 0; t < E->nks; t++) {
        const double r = pol_radj(E, id, E->ks[t]);
        if (r > best) best = r;
    }
    return isfinite(best) ? best : 0.0;
}
/* ball_occludes_distance (:474-490), both sides fused as GCC builds them */
static int pol_ball_occ(const or_edge_policy* E, float alpha, uint32_t base, uint32_t sid, float dbase, float dsc) {
    if (!E->ball || E->nks == 0) return 0;
    int any = 0, all = 1;
    for (uint32_t t = 0; t < E->nks; t++) {
        const double lhs = fma(-E->lambda, pol_radj(E, sid, E->ks[t]), (double)alpha * (double)dsc);
        const double rhs = fma(-E->lambda, pol_radj(E, base, E->ks[t]), (double)dbase);
        const int occ = lhs < rhs;
        any = any || occ;
        all = all && occ;
        if (!E->occlude_all && any) return 1;
        if (E->occlude_all && !all) return 0;
    }
    return E->occlude_all ? all : any;
}
/* lsg_distance2 (:832-838) */
static double pol_lsg(const or_edge_policy* E, uint32_t a, uint32_t b, double raw) {
    if (!E->lsg || E->lsg_alpha == 0.0) return raw;
    const double mp = (double)E->mu[a] * (double)E->mu[b];
    const double den = pow(mp, E->lsg_alpha);
    if (!isfinite(den) || den <= 0.0) return raw;
    return raw / den;
}

typedef struct { double k; uint32_t id; float d; } pol_key;
static int pol_key_cmp(const void* a, const void* b) {
    const pol_key* x = (const pol_key*)a;
    const pol_key* y = (const pol_key*)b;
    if (x->k < y->k) return -1;
    if (x->k > y->k) return 1;
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

/* rng_extend (:267-281) over (d, id) */
static uint32_t pol_rng(const float* x, const float* nrm, uint32_t d, float alpha, const float* pd, const uint32_t* pi,
                        uint32_t m, uint32_t cap, uint32_t* ids, uint32_t deg) {
    for (uint32_t t = 0; t < m; t++) {
        if (deg >= cap) break;
        const uint32_t cid = pi[t];
        if (pol_has(ids, deg, cid)) continue;
        int occ = 0;
        for (uint32_t s = 0; s < deg && !occ; s++) occ = alpha * pol_d2(x, nrm, d, ids[s], cid) < pd[t];
        if (!occ) ids[deg++] = cid;
    }
    return deg;
}
/* rng_extend_ball (:505-519) over the ball order, occlusion by ball_occludes */
static uint32_t pol_rng_ball(const float* x, const float* nrm, uint32_t d, float alpha, const or_edge_policy* E,
                             uint32_t base, const pol_key* bp, uint32_t m, uint32_t cap, uint32_t* ids, uint32_t deg) {
    for (uint32_t t = 0; t < m; t++) {
        if (deg >= cap) break;
        const uint32_t cid = bp[t].id;
        if (pol_has(ids, deg, cid)) continue;
        int occ = 0;
        for (uint32_t s = 0; s < deg && !occ; s++)
            occ = pol_ball_occ(E, alpha, base, ids[s], bp[t].d, pol_d2(x, nrm, d, ids[s], cid));
        if (!occ) ids[deg++] = cid;
    }
    return deg;
}
/* rng_extend_lsg (:853-871) over the lsg order */
static uint32_t pol_rng_lsg(const float* x, const float* nrm, uint32_t d, float alpha, const or_edge_policy* E,
                            const pol_key* lp, uint32_t m, uint32_t cap, uint32_t* ids, uint32_t deg) {
    for (uint32_t t = 0; t < m; t++) {
        if (deg >= cap) break;
        const uint32_t cid = lp[t].id;
        if (pol_has(ids, deg, cid)) continue;
        int occ = 0;
        for (uint32_t s = 0; s < deg && !occ; s++) {
            const double dsc = pol_lsg(E, ids[s], cid, (double)pol_d2(x, nrm, d, ids[s], cid));
            occ = (double)alpha * dsc < lp[t].k;
        }
        if (!occ) ids[deg++] = cid;
    }
    return deg;
}
/* add_residual_cover_edges (:441-472) */
static uint32_t pol_cover(const or_edge_policy* E, uint32_t n, uint32_t base, const float* pd, const uint32_t* pi,
                          uint32_t m, uint32_t M, uint32_t* ids, uint32_t deg) {
    if (!E->resid_cover || deg >= M) return deg;
    const uint32_t total = E->budget < M - deg ? E->budget : M - deg;
    if (total == 0 || E->nks == 0) return deg;
    uint32_t per_k = (total + E->nks - 1) / E->nks;
    if (per_k < 1) per_k = 1;
    uint32_t added = 0;
    pol_key* sc = (pol_key*)malloc(sizeof(pol_key) * (m ? m : 1));
    for (uint32_t t = 0; t < E->nks; t++) {
        if (added >= total || deg >= M) break;
        const int k = E->ks[t];
        uint32_t c = 0;
        for (uint32_t u = 0; u < m; u++) {
            const uint32_t cid = pi[u];
            if (cid == UINT32_MAX || cid >= n || cid == base || pol_has(ids, deg, cid)) continue;
            const double r = (double)E->resid[(size_t)cid * E->nk + (size_t)(k - 1)];
            const double risk = r > 0.0 ? r : 0.0;
            if (risk <= 0.0) continue;
            sc[c].k = fma(-(E->lambda * E->scale2), risk, (double)pd[u]);
            sc[c].id = cid;
            c++;
        }
        qsort(sc, c, sizeof(pol_key), pol_key_cmp);
        uint32_t ak = 0;
        for (uint32_t u = 0; u < c; u++) {
            if (deg >= M || added >= total || ak >= per_k) break;
            if (pol_has(ids, deg, sc[u].id)) continue;
            ids[deg++] = sc[u].id;
            ak++;
            added++;
        }
    }
    free(sc);
    return deg;
}

/* optimize_verification_constrained_bubble (:521-723): with the plain
 * prune's ids selected, greedy swaps against the top alternatives in the ball
 * order (or the whole pool) that keep every target some selected edge
 * verifies (alpha d2(s, t) < d2(base, t), or s == t) verified, and raise
 * the number of targets a selected edge ball-occludes; ties by the power key
 * gain.  ids (deg of them) are replaced in place. */
static void pol_bubble(const float* x, const float* nrm, uint32_t d, float alpha, const or_edge_policy* E,
                       uint32_t base, const float* pd, const uint32_t* pi, uint32_t w, const pol_key* bp,
                       uint32_t* ids, uint32_t deg) {
    if (!E->bubble_auto || deg == 0 || w == 0) return;
    uint32_t* U = (uint32_t*)malloc(sizeof(uint32_t) * (deg + w));
    uint32_t nu = 0;
    for (uint32_t t = 0; t < deg; t++) U[nu++] = ids[t];
    const uint32_t limit = E->bubble_full ? w : E->auto_candidates;
    uint32_t alts = 0;
    for (uint32_t t = 0; t < w; t++) {
        if (pol_has(U, nu, bp[t].id)) continue;
        U[nu++] = bp[t].id;
        if (++alts >= limit) break;
    }
    if (alts == 0) { free(U); return; }
    /* targets: the whole pool, or the universe's own entries, by (d, id) */
    uint32_t nt = 0;
    float* td = (float*)malloc(sizeof(float) * (w + nu));
    uint32_t* tid = (uint32_t*)malloc(sizeof(uint32_t) * (w + nu));
    if (E->bubble_full) {
        for (uint32_t t = 0; t < w; t++) { td[nt] = pd[t]; tid[nt] = pi[t]; nt++; }
    } else {
        for (uint32_t u = 0; u < nu; u++)
            for (uint32_t t = 0; t < w; t++)
                if (pi[t] == U[u]) { td[nt] = pd[t]; tid[nt] = pi[t]; nt++; break; }
        or_sort_pairs(td, tid, nt);
    }
    uint8_t* ver = (uint8_t*)calloc((size_t)nu * nt, 1);
    uint8_t* rou = (uint8_t*)calloc((size_t)nu * nt, 1);
    double* pk = (double*)malloc(sizeof(double) * nu);
    for (uint32_t u = 0; u < nu; u++) {
        float db = 0.0f;
        for (uint32_t t = 0; t < w; t++) if (pi[t] == U[u]) { db = pd[t]; break; }
        pk[u] = fma(-E->lambda, pol_sort_radj(E, U[u]), (double)db);
        for (uint32_t t = 0; t < nt; t++) {
            const float dsc = U[u] != tid[t] ? pol_d2(x, nrm, d, U[u], tid[t]) : 0.0f;
            ver[(size_t)u * nt + t] = U[u] == tid[t] || alpha * dsc < td[t];
            rou[(size_t)u * nt + t] = (uint8_t)pol_ball_occ(E, alpha, base, U[u], td[t], dsc);
        }
    }
    uint32_t* sel = (uint32_t*)malloc(sizeof(uint32_t) * deg);
    uint8_t* is_sel = (uint8_t*)calloc(nu, 1);
    uint32_t ns = 0;
    for (uint32_t t = 0; t < deg; t++)
        for (uint32_t u = 0; u < nu; u++)
            if (U[u] == ids[t]) { sel[ns++] = u; is_sel[u] = 1; break; }
    int* vc = (int*)calloc(nt, sizeof(int));
    int* rc = (int*)calloc(nt, sizeof(int));
    uint8_t* req = (uint8_t*)calloc(nt, 1);
    for (uint32_t s2 = 0; s2 < ns; s2++)
        for (uint32_t t = 0; t < nt; t++) { vc[t] += ver[(size_t)sel[s2] * nt + t]; rc[t] += rou[(size_t)sel[s2] * nt + t]; }
    for (uint32_t t = 0; t < nt; t++) req[t] = vc[t] > 0;
    uint32_t swaps = 0;
    const uint32_t max_swaps = E->auto_max_swaps < alts ? E->auto_max_swaps : alts;
    while (swaps < max_swaps) {
        long best_rg = LONG_MIN;
        double best_pg = -INFINITY;
        uint32_t bq = UINT32_MAX, bs = UINT32_MAX;
        for (uint32_t q = 0; q < nu; q++) {
            if (is_sel[q]) continue;
            long newly = 0;
            for (uint32_t t = 0; t < nt; t++) newly += rc[t] == 0 && rou[(size_t)q * nt + t];
            for (uint32_t s2 = 0; s2 < ns; s2++) {
                const uint32_t old = sel[s2];
                int feasible = 1;
                long lost = 0;
                for (uint32_t t = 0; t < nt; t++) {
                    const int ov = ver[(size_t)old * nt + t], qv = ver[(size_t)q * nt + t];
                    if (req[t] && ov && vc[t] == 1 && !qv) { feasible = 0; break; }
                    const int orr = rou[(size_t)old * nt + t], qr = rou[(size_t)q * nt + t];
                    if (orr && rc[t] == 1 && !qr) lost++;
                }
                if (!feasible) continue;
                const long rg = newly - lost;
                const double pg = pk[old] - pk[q];
                if (rg > best_rg || (rg == best_rg && pg > best_pg)) { best_rg = rg; best_pg = pg; bq = q; bs = s2; }
            }
        }
        if (bq == UINT32_MAX || best_rg < 0 || (best_rg == 0 && best_pg <= 1e-12)) break;
        const uint32_t old = sel[bs];
        for (uint32_t t = 0; t < nt; t++) {
            vc[t] += (int)ver[(size_t)bq * nt + t] - (int)ver[(size_t)old * nt + t];
            rc[t] += (int)rou[(size_t)bq * nt + t] - (int)rou[(size_t)old * nt + t];
        }
        is_sel[old] = 0;
        is_sel[bq] = 1;
        sel[bs] = bq;
        swaps++;
    }
    for (uint32_t s2 = 0; s2 < ns; s2++) ids[s2] = U[sel[s2]];
    free(U); free(td); free(tid); free(ver); free(rou); free(pk); free(sel); free(is_sel); free(vc); free(rc); free(req);
}

/* prune_pool_to_ids (:772-829) under an edge policy; the sort_unique step as
 * in or_prune.  Policies whose configuration left every flag off prune like
 * or_prune. */
uint32_t or_prune_policy(const float* x, const float* nrm, uint32_t d, uint32_t base, uint32_t n, float* pool_d,
                         uint32_t* pool_id, uint32_t m, uint32_t M, float alpha, const or_edge_policy* E,
                         uint32_t* out) {
    /* sort_unique_candidates (:225-248) */
    uint32_t w = 0;
    for (uint32_t t = 0; t < m; t++) {
        uint32_t id = pool_id[t];
        if (id == UINT32_MAX || id >= n || id == base) continue;
        pool_id[w] = id; pool_d[w] = pool_d[t]; w++;
    }
    {
        or_prop* P = (or_prop*)malloc(sizeof(or_prop) * (w ? w : 1));
        for (uint32_t t = 0; t < w; t++) { P[t].row = pool_id[t]; P[t].id = 0; P[t].d = pool_d[t]; }
        qsort(P, w, sizeof(or_prop), prop_cmp);
        uint32_t u = 0;
        for (uint32_t r = 0; r < w;) {
            uint32_t e = r + 1;
            while (e < w && P[e].row == P[r].row) e++;
            pool_id[u] = P[r].row; pool_d[u] = P[r].d; u++;
            r = e;
        }
        free(P);
        w = u;
    }
    or_sort_pairs(pool_d, pool_id, w);
    uint32_t deg = 0;
    pol_key* lp = (pol_key*)malloc(sizeof(pol_key) * (w ? w : 1));
    pol_key* bp = (pol_key*)malloc(sizeof(pol_key) * (w ? w : 1));
    /* lsg_sort_candidates (:840-851); ball_sort_candidates (:494-503, key
     * d - lambda * max_k radj, fused) */
    for (uint32_t t = 0; t < w; t++) {
        lp[t].k = pol_lsg(E, base, pool_id[t], (double)pool_d[t]);
        lp[t].id = pool_id[t];
        lp[t].d = pool_d[t];
        bp[t].k = E->ball ? fma(-E->lambda, pol_sort_radj(E, pool_id[t]), (double)pool_d[t]) : (double)t;
        bp[t].id = pool_id[t];
        bp[t].d = pool_d[t];
    }
    qsort(lp, w, sizeof(pol_key), pol_key_cmp);
    if (E->ball) qsort(bp, w, sizeof(pol_key), pol_key_cmp);
    const uint32_t bud = E->budget < M ? E->budget : M;
    if (E->bubble_auto) {
        deg = pol_rng(x, nrm, d, alpha, pool_d, pool_id, w, M, out, deg);
        pol_bubble(x, nrm, d, alpha, E, base, pool_d, pool_id, w, bp, out, deg);
    } else if (E->lsg && E->ball && E->lsg_alpha != 0.0) {
        uint32_t cap = M - bud;
        if (cap == 0) cap = M;
        deg = pol_rng_lsg(x, nrm, d, alpha, E, lp, w, cap, out, deg);
        deg = pol_rng_ball(x, nrm, d, alpha, E, base, bp, w, M, out, deg);
        deg = pol_rng_lsg(x, nrm, d, alpha, E, lp, w, M, out, deg);
    } else if (E->lsg && E->lsg_alpha != 0.0) {
        deg = pol_rng_lsg(x, nrm, d, alpha, E, lp, w, M, out, deg);
    } else if (E->ball) {
        if (E->ball_pure) {
            deg = pol_rng_ball(x, nrm, d, alpha, E, base, bp, w, M, out, deg);
            deg = pol_rng(x, nrm, d, alpha, pool_d, pool_id, w, M, out, deg);
        } else {
            uint32_t cap = M - bud;
            if (cap == 0) cap = M;
            deg = pol_rng(x, nrm, d, alpha, pool_d, pool_id, w, cap, out, deg);
            deg = pol_rng_ball(x, nrm, d, alpha, E, base, bp, w, M, out, deg);
            deg = pol_rng(x, nrm, d, alpha, pool_d, pool_id, w, M, out, deg);
        }
    } else {
        const uint32_t cb = E->resid_cover ? bud : 0;
        uint32_t cap = M - cb;
        if (cap == 0) cap = M;
        deg = pol_rng(x, nrm, d, alpha, pool_d, pool_id, w, cap, out, deg);
        deg = pol_cover(E, n, base, pool_d, pool_id, w, M, out, deg);
        deg = pol_rng(x, nrm, d, alpha, pool_d, pool_id, w, M, out, deg);
    }
    free(lp); free(bp);
    return deg;
}
