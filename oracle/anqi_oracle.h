/*
 * anqi_oracle.h -- CPU restatement of the ANQI approximate reverse-kNN path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the HIP library, the
 * command-line tools, bench.py's timed region) may link, load or call this
 * code.  tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it
 * as the checker.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Pinning status of each function (which
 * golden vector / known-answer test / invariant checks it) is listed in the
 * header comment of anqi_oracle.c.
 */
#ifndef ANQI_ORACLE_H
#define ANQI_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FP32 distance contract (include/distance.h) ------------------------ */
float or_ip(const float* x, const float* y, uint32_t d);      /* distance.h:22-49, 89-115 */
float or_l2(const float* x, const float* y, uint32_t d);      /* distance.h:58-86 */
float or_norm(const float* x, uint32_t d);                    /* distance.h:143-161 */
/* nndindex::d2, include/index.h:172-176: na + nb - 2*ip (contracted). */
float or_d2(const float* x, const float* y, float nx, float ny, uint32_t d);
void or_norms(const float* x, uint64_t n, uint32_t d, float* out);

/* ---- exact top-K and RkNN ground truth (tools/dataset_gt.cpp) ------------ */
/* dataset_gt.cpp:206-245.  ids/keys: n*K, rows sorted by (key, id). */
int or_exact_topk(const float* base, uint32_t n, uint32_t d, uint32_t K,
                  uint32_t* ids, float* keys, int threads);
/* dataset_gt.cpp:283-313.  Returns total hits; off has nq+1 entries.  Call
 * with ids==NULL to size, then again with buffers of off[nq] entries. */
uint64_t or_rknn_membership(const float* base, uint32_t n, const float* query,
                            uint32_t nq, uint32_t d, const float* rk2,
                            uint32_t* off, uint32_t* ids, float* keys, int threads);
/* The same two with the metric: key = -<x, y> for "ip" (dataset_gt.cpp:225, 292). */
int or_exact_topk_metric(const float* base, uint32_t n, uint32_t d, uint32_t K, int ip_metric,
                         uint32_t* ids, float* keys, int threads);
uint64_t or_rknn_membership_metric(const float* base, uint32_t n, const float* query,
                                   uint32_t nq, uint32_t d, int ip_metric, const float* rk2,
                                   uint32_t* off, uint32_t* ids, float* keys, int threads);

/* ---- bounded max-heap rows (include/heap.h) ------------------------------ */
/* HeapList::push_nolock / check_push(update_reverse=false), heap.h:83-186.
 * idx/val: one row of capacity cap (sentinel UINT32_MAX / FLT_MAX). */
int or_heap_push(uint32_t* idx, float* val, uint32_t cap, uint32_t node, float value);

/* ---- sorted candidate queue (include/neighbor.h) ------------------------- */
typedef struct {
    float* dist;
    uint32_t* pid;   /* bit31 = expanded */
    uint32_t size, cap, cur;
} or_queue;
void or_queue_init(or_queue* q, float* dist, uint32_t* pid, uint32_t cap);
int or_queue_insert(or_queue* q, uint32_t id, float dist);          /* neighbor.h:61-102 */
uint32_t or_queue_pop_closest(or_queue* q);                         /* neighbor.h:105-111 */

/* ---- RP tree (include/rptrees.h) ----------------------------------------- */
/* A tree in flat form: node k is internal if left[k] != UINT32_MAX.
 * hyper: nodes*d floats (only internal nodes meaningful); leaf_off/leaf_ids
 * give each leaf's indices in stored order.  root = nodes-1 (post-order). */
typedef struct {
    uint32_t nodes, d;
    const uint64_t* left;
    const uint64_t* right;
    const float* offset;
    const float* hyper;
    const uint64_t* idx_off;   /* nodes+1 */
    const uint32_t* idx;       /* leaf indices */
} or_tree;
/* rptree::get_leave_mediod, rptrees.h:108-123 (ties |margin|<EPS go left). */
uint32_t or_tree_entry(const or_tree* t, const float* q);

/* generate_random_integer (include/utils.h:169-173): one thread_local
 * std::mt19937 per thread, seeded by the first call's seed only, drawn through
 * libstdc++'s uniform_int_distribution<uint64_t> (GCC 11
 * bits/uniform_int_dist.h: a 32-bit engine takes Lemire's nearly-divisionless
 * branch).  or_mt_* is MT19937 (Matsumoto & Nishimura 1998). */
typedef struct { uint32_t x[624]; uint32_t i; } or_mt;
void or_mt_seed(or_mt* m, uint64_t seed);
uint32_t or_mt_next(or_mt* m);
uint32_t or_uniform_below(or_mt* m, uint32_t range);   /* uniform in [0, range) */
void or_mt_stream(uint64_t seed, uint32_t* out, uint64_t count);
/* build_rptree (rptrees.h:28-163) on one fresh generator seeded with seed: the
 * whole tree draws from one stream in depth-first order.  Writes nodes in the
 * reference's post-order (children before parent, root last; leaves have
 * left == right == UINT32_MAX); idx_off/idx hold each leaf's indices in stored
 * order.  node_cap bounds nodes, n bounds idx.  Returns the node count (0 if
 * node_cap is too small); *draws = engine outputs consumed. */
uint32_t or_rp_tree_build(const float* data, uint32_t n, uint32_t d, uint32_t leaf,
                          uint32_t max_depth, uint64_t seed, uint32_t node_cap,
                          uint64_t* left, uint64_t* right, float* offset, float* hyper,
                          uint64_t* idx_off, uint32_t* idx, uint64_t* draws);

/* ---- range search (include/index.h:1535-1652, CENTER mode, radj==NULL) --- */
/* lifted: n*dl, deg: n, links: n*M.  Writes up to cap candidate ids in
 * discovery order; returns count (may exceed cap: truncated).  ndist out. */
uint32_t or_range_search(const float* lifted, uint32_t n, uint32_t dl,
                         const uint32_t* deg, const uint32_t* links, uint32_t M,
                         uint32_t entry, const float* q, double R2, uint32_t L,
                         uint32_t* out, uint32_t cap, uint64_t* ndist);

/* Query transform, src/rknn_query.cpp:2719-2730 (anyk_lift, L2). */
void or_lift_query(const float* q, uint32_t d, double scale, float* qb);

/* Postfilter key, src/rknn_query.cpp:2555-2561. */
float or_recheck_key(const float* o, const float* q, uint32_t d);
/* Inner-product metric ("ip"): postfilter key -<o, q> (:2555-2561); the
 * lifted query and R2 for both metrics (:2719-2730; ip_metric 0: L2); the
 * nndindex "ip" distance 1 - <x, y> (include/distance.h:170-175). */
float or_recheck_key_ip(const float* o, const float* q, uint32_t d);
double or_query_lift(const float* q, uint32_t d, double scale, double mstar2, int ip_metric, float* qb);
float or_ipdist(const float* x, const float* y, uint32_t d);

/* ---- LRQ verifier (src/rknn_query.cpp:603-839), FP64 on the host --------- */
uint32_t or_lrq_encode_value(double v, const float* centers, uint32_t code_size); /* :482-513 floor */
void or_lrq_encode(const float* cutoffs, uint64_t n, uint32_t nk, uint32_t rankm,
                   const float* mu, const float* basis, const float* codebook,
                   uint32_t code_size, float* coeff, uint8_t* codes);            /* :679-720 */
void or_lrq_materialize(const float* coeff, const uint8_t* codes, uint64_t n,
                        uint32_t nk, uint32_t rankm, const float* mu,
                        const float* basis, const float* codebook,
                        uint32_t code_size, uint32_t k_one_based, double slack,
                        float* thr);                                              /* :722-753 */
void or_lrq_materialize_metric(const float* coeff, const uint8_t* codes, uint64_t n,
                               uint32_t nk, uint32_t rankm, const float* mu,
                               const float* basis, const float* codebook,
                               uint32_t code_size, uint32_t k_one_based, double slack,
                               int ip_metric, float* thr);                        /* :722-753, :751 */

/* ---- lift (src/rknn_query.cpp:2203-2236), FP64 on the host --------------- */
/* Writes n*(d+2) lifted floats; returns M*^2 and scale through pointers.
 * sum of squared norms is accumulated in row order (the reference uses an
 * OpenMP reduction; identical with one thread). */
void or_lift(const float* base, uint64_t n, uint32_t d, const float* rk_horizon,
             float* lifted, double* mstar2, double* scale);
void or_lift_metric(const float* base, uint64_t n, uint32_t d, const float* rk_horizon, int ip_metric,
                    float* lifted, double* mstar2, double* scale);   /* :2227 for "ip" */

/* ---- graph build pieces (include/index.h) -------------------------------- */
/* Warm start, index.h:1204-1218: push width AKNN ids with lifted d2 into rows
 * of capacity cap (heap arrays n*cap, initialised here). */
void or_warm_start(const float* x, const float* nrm, uint32_t n, uint32_t d,
                   const uint32_t* ids, uint32_t width, uint32_t cap,
                   uint32_t* hidx, float* hval);
/* nndescent_refine, index.h:1468-1487, single-threaded order (u ascending,
 * updates visible immediately).  Returns passes run; upd per pass in upd. */
int or_nndescent_seq(const float* x, const float* nrm, uint32_t n, uint32_t d,
                     uint32_t cap, uint32_t* hidx, float* hval, int iters,
                     int sample, int64_t* upd);
/* Deterministic snapshot schedule used by the GPU: each pass reads every row's
 * sample at pass start, then applies each row's proposals in ascending
 * (dist, id) order.  Same heap semantics. */
int or_nndescent_snapshot(const float* x, const float* nrm, uint32_t n, uint32_t d,
                          uint32_t cap, uint32_t* hidx, float* hval, int iters,
                          int sample, int64_t* upd);
/* HeapList::push_reverse, heap.h:190-207, in row order (1 thread). */
void or_reverse_csr(const uint32_t* hidx, uint32_t n, uint32_t cap,
                    uint32_t* roff, uint32_t* rids);
/* gather_candidates, index.h:1491-1525.  pool_d/pool_id receive the pool in
 * push order (capacity pool_max); returns pool size. */
uint32_t or_gather(const float* x, const float* nrm, uint32_t n, uint32_t d,
                   const uint32_t* hidx, const float* hval, uint32_t cap,
                   const uint32_t* roff, const uint32_t* rids, uint32_t i,
                   uint32_t explore, uint32_t pool_cap, float* pool_d,
                   uint32_t* pool_id, uint32_t pool_max);
/* build_approx_knn tail, index.h:1348-1366: sort pool by (dist,id), drop self,
 * keep k.  Returns kept count. */
uint32_t or_pool_topk(uint32_t self, const float* pool_d, const uint32_t* pool_id,
                      uint32_t m, uint32_t k, uint32_t* ids, float* keys);
/* prune_pool_to_ids default branch, index.h:772-829 with sort_unique_candidates
 * (:225-248) and rng_extend (:267-281). Returns degree. */
uint32_t or_prune(const float* x, const float* nrm, uint32_t d, uint32_t base,
                  uint32_t n, float* pool_d, uint32_t* pool_id, uint32_t m,
                  uint32_t M, float alpha, uint32_t* out);
/* commit_edges, index.h:907-930.  O: n*M with counts ocnt; writes links/deg. */
void or_commit(const float* x, const float* nrm, uint32_t n, uint32_t d,
               const uint32_t* O, const uint32_t* ocnt, uint32_t M, float alpha,
               uint32_t* links, uint32_t* deg);

/* ---- Vamana refinement (include/index.h:831-1065, 1262-1282) -------------- */
/* greedy_collect / greedy_collect_from_entry (index.h:831-880): pool in visit
 * order (row i itself excluded); returns its size (truncated at pool_max). */
uint32_t or_vamana_collect(const float* x, const float* nrm, uint32_t n, uint32_t d,
                           const uint32_t* deg, const uint32_t* links, uint32_t M,
                           uint32_t i, uint32_t entry, uint32_t Lb,
                           float* pool_d, uint32_t* pool_id, uint32_t pool_max);
/* One pass of ANQI_VAMANA_STYLE=global (index.h:1267-1279); entries[i] = the
 * tree entry of lifted row i; links/deg updated in place. */
void or_vamana_global_pass(const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* entries,
                           const uint32_t* hidx, uint32_t cap, uint32_t Lb, uint32_t M, float alpha,
                           uint32_t* links, uint32_t* deg);
/* ANQI_VAMANA_STYLE=batch (vamana_batch_refine, index.h:932-1063). */
void or_vamana_batch(const float* x, const float* nrm, uint32_t n, uint32_t d, const uint32_t* hidx, uint32_t cap,
                     uint32_t Lb, int passes, double batch_frac, uint64_t max_batch_cap, double batch_base,
                     float alpha, float early_alpha, int use_early, uint32_t entry, uint64_t seed, uint32_t M,
                     uint32_t* links, uint32_t* deg);
/* choose_vamana_entry "medoid"/"centroid" (index.h:190-222). */
uint32_t or_vamana_medoid(const float* x, uint32_t n, uint32_t d, uint64_t sample);
/* std::mt19937_64 and libstdc++'s std::shuffle over it (GCC 11). */
typedef struct { uint64_t x[312]; uint32_t i; } or_mt64;
void or_mt64_seed(or_mt64* m, uint64_t seed);
uint64_t or_mt64_next(or_mt64* m);
void or_shuffle_u32(uint32_t* a, uint32_t n, or_mt64* g);

/* ---- edge policies (include/index.h:283-829) ------------------------------- */
/* The configured state of configure_edge_residual_cover (:305-438): which
 * branch prune_pool_to_ids takes and its inputs.  resid: n*nk residuals;
 * ks: 1-based ranks (<= nk); mu: n lsg weights. */
typedef struct {
    int resid_cover, ball, ball_pure, occlude_all, lsg;
    double lsg_alpha;
    const float* mu;
    const float* resid;
    uint32_t nk;
    const int32_t* ks;
    uint32_t nks;
    uint32_t budget;
    double lambda, scale2;
    int bubble_auto, bubble_full;              /* bubble_auto, bubble_auto_full */
    uint32_t auto_candidates, auto_max_swaps;  /* ANQI_EDGE_AUTO_CANDIDATES / _MAX_SWAPS */
} or_edge_policy;
uint32_t or_prune_policy(const float* x, const float* nrm, uint32_t d, uint32_t base, uint32_t n, float* pool_d,
                         uint32_t* pool_id, uint32_t m, uint32_t M, float alpha, const or_edge_policy* E,
                         uint32_t* out);
/* The policy or_commit and or_vamana_* prune with from now on (NULL: plain). */
void or_set_edge_policy(const or_edge_policy* E);

/* ---- small helpers shared by tests -------------------------------------- */
void or_sort_pairs(float* d, uint32_t* id, uint32_t m);   /* (dist, id) ascending */

#ifdef __cplusplus
}
#endif
#endif
