/*
 * anqi.h -- C ABI of the MI355X ANQI reverse-kNN library (libanqi_hip.so).
 *
 * The reference exposes no library API: its boundary is the processes
 * build_aknn_source and rknn_query (argv, ANQI_* environment, on-disk files,
 * stdout lines, exit codes).  This ABI is the layer the drop-in command-line
 * tools (anqi-implementation_amd/cli/) call; each entry point names the
 * reference code it replaces.  Conventions:
 *   - plain pointers and sizes; "device" arguments are HIP device pointers,
 *     "host" arguments are host pointers;
 *   - every call returns an int status (ANQI_OK or one of the codes below,
 *     which the tools map to the reference exit codes) and never throws;
 *     anqi_last_error() describes the last failure;
 *   - work is enqueued on the context's stream; calls that must read results
 *     on the host (counts, overflow checks) synchronise internally;
 *   - a context is used by one host thread at a time.
 */
#ifndef ANQI_H
#define ANQI_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ANQI_OK 0
#define ANQI_E_USAGE 2     /* bad arguments (reference exit 2) */
#define ANQI_E_INPUT 3     /* bad input / format (exit 3) */
#define ANQI_E_ROWS 4      /* row-count mismatch (exit 4) */
#define ANQI_E_INVALID 5   /* invalid output row (exit 5) */
#define ANQI_E_IO 6        /* write failure (exit 6) */
#define ANQI_E_DEVICE 7    /* HIP / device failure */

/* Metrics (the reference's "l2" and "ip", include/distance.h:178-186;
 * src/rknn_query.cpp:1611-1617) */
#define ANQI_METRIC_L2 0
#define ANQI_METRIC_IP 1

typedef struct anqi_ctx anqi_ctx;

/* ---- context ------------------------------------------------------------- */
int anqi_ctx_create(int device, anqi_ctx** out);
void anqi_ctx_destroy(anqi_ctx* ctx);
const char* anqi_last_error(const anqi_ctx* ctx);
int anqi_ctx_sync(anqi_ctx* ctx);
/* Enqueue on an external hipStream_t (e.g. the caller's); NULL selects the
 * default (null) stream.  anqi_ctx_reset_stream restores the context's own. */
int anqi_ctx_set_stream(anqi_ctx* ctx, void* stream);
int anqi_ctx_reset_stream(anqi_ctx* ctx);
/* Timing of the range-search kernel launch itself (HIP events recorded on
 * the context's stream around it, excluding the batch's entry descent,
 * ordering and result copies): enable, then read the last call's
 * milliseconds.  No counterpart in the reference (measurement support). */
int anqi_ctx_time_kernels(anqi_ctx* ctx, int enable);
double anqi_ctx_last_kernel_ms(const anqi_ctx* ctx);
/* Queries of the last anqi_rknn_search* call on the context that were rerun
 * because they outgrew their LDS work areas: with the 16K-entry retry areas
 * (*retry) and then in the global-memory tier (*global; with the per-query
 * visited bitmap and FIFO in HBM, as unbounded as the reference's flood,
 * include/index.h:33-52).  Either pointer may be NULL.  No counterpart in
 * the reference (measurement support). */
int anqi_ctx_redo_counts(const anqi_ctx* ctx, uint32_t* retry, uint32_t* global);
void* anqi_ctx_stream(anqi_ctx* ctx);
/* Work areas and grid of the persistent search (mode 1 of anqi_rknn_search
 * and the batcher); anqi_search_autotune measures candidates.  Zero fields
 * keep the defaults. */
typedef struct {
    uint32_t waves_per_cu;    /* persistent grid = waves_per_cu x CUs (0: occupancy) */
    uint32_t visited_slots;   /* LDS visited-set slots, rounded up to a multiple of 4 (0: from L) */
    uint32_t result_cap;      /* LDS candidate FIFO entries (0: 1024); later
                                 entries go to the query's output row */
} anqi_search_tuning;
/* Tuning used by later mode-1 searches on this context (NULL: defaults). */
int anqi_ctx_set_search_tuning(anqi_ctx* ctx, const anqi_search_tuning* tune);
int anqi_dmalloc(anqi_ctx* ctx, size_t bytes, void** out);
int anqi_dfree(anqi_ctx* ctx, void* p);
int anqi_h2d(anqi_ctx* ctx, void* dst, const void* src, size_t bytes);
int anqi_d2h(anqi_ctx* ctx, void* dst, const void* src, size_t bytes);

/* ---- FP32 distance contract (include/distance.h:22-161, index.h:172-176) -- */
/* out[i] = <x_i, x_i>.  device x (n*d), out (n). */
int anqi_norms(anqi_ctx* ctx, const float* x, uint64_t n, uint32_t d, float* out);
/* out[t] = nrm[a] + nrm[b] - 2 <x_a, x_b>.  all device. */
int anqi_pair_d2(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t d,
                 const uint32_t* a, const uint32_t* b, uint64_t m, float* out);

/* ---- exact ground truth (tools/dataset_gt.cpp:206-313) ------------------- */
/* Self-excluded top-K by (key, id).  precision 0: FP32 tiles; 1: bf16 MFMA
 * tiles with exact FP32 re-scoring (same output).  device x, ids, keys (n*K). */
int anqi_exact_knn(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, uint32_t K,
                   int precision, uint32_t* ids, float* keys);
/* Reverse-kNN membership key <= rk2[j].  host off (nq+1) and *total out;
 * ids/keys are device buffers of *total entries, each row sorted by (key, id).
 * Call with ids == NULL to size. */
int anqi_rknn_membership(anqi_ctx* ctx, const float* base, uint32_t n, const float* q,
                         uint32_t nq, uint32_t d, const float* rk2, uint32_t* off,
                         uint64_t* total, uint32_t* ids, float* keys);
/* The same two with the metric: ANQI_METRIC_IP orders and tests by the key
 * -<x, y> (dataset_gt.cpp:225, 292, "--metric ip"); the first two are the
 * L2 forms. */
int anqi_exact_knn_metric(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, uint32_t K,
                          int precision, int metric, uint32_t* ids, float* keys);
int anqi_rknn_membership_metric(anqi_ctx* ctx, const float* base, uint32_t n, const float* q,
                                uint32_t nq, uint32_t d, int metric, const float* rk2, uint32_t* off,
                                uint64_t* total, uint32_t* ids, float* keys);

/* ---- AKNN source (build_aknn_source, include/index.h:1300-1461) ----------- */
typedef struct {
    uint32_t K;          /* 100 */
    uint32_t leaf;       /* 500 */
    uint32_t trees;      /* 64 */
    uint32_t ef;         /* explore range, 600 */
    uint32_t ef_alpha;   /* pool cap = max(K, ef_alpha * ef), 8 */
    uint32_t max_depth;  /* 100 */
    uint64_t seed;       /* 10 */
    int precision;       /* 0 FP32 leaf tiles, 1 bf16 MFMA filtered, 2 FP8-e4m3
                            MFMA filtered (trial); every mode gives the same lists */
    int metric;          /* ANQI_METRIC_L2 (0) or ANQI_METRIC_IP: "ip" splits on
                            unit-difference hyperplanes (rptrees.h:41-52) and ranks
                            by 1 - <x, y> (distance.h:170-175); its leaf pairs run
                            in the FP32 kernel whatever the precision */
} anqi_aknn_params;
typedef struct {
    double forest_s, leaf_s, gather_s, total_s;
    uint64_t leaf_pairs;
    uint32_t leaves, max_leaf, short_rows, overflow_rows;
    uint64_t leaf_survivors;  /* precision 1/2: ordered pairs the MFMA filter let through */
    uint64_t gather_d2, gather_adj;  /* gather work: distances computed, adjacency ids read */
} anqi_aknn_stats;
/* device x (n*d); device ids/keys (n*K) rows sorted by (dist, id), self
 * excluded; keys are the raw distances (d2, or 1 - <x, y> for "ip"; the
 * tool writes max(0, d2), or the distance minus 1 for "ip", into the .rec). */
int anqi_aknn_build(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d,
                    const anqi_aknn_params* p, uint32_t* ids, float* keys,
                    anqi_aknn_stats* stats);

/* ---- LRQ verifier (src/rknn_query.cpp:603-753, policy floor, u8) ---------- */
/* device everything.  codes row-major (n*nk, file order); codes_kmajor
 * (nk*n) is the device layout materialize reads (may be NULL). */
int anqi_lrq_encode(anqi_ctx* ctx, const float* cutoffs, uint64_t n, uint32_t nk,
                    uint32_t rankm, const float* mu, const float* basis,
                    const float* codebook, uint32_t code_size, float* coeff,
                    uint8_t* codes, uint8_t* codes_kmajor);
int anqi_transpose_codes(anqi_ctx* ctx, const uint8_t* codes, uint64_t n, uint32_t nk,
                         uint8_t* codes_kmajor);
/* thr[i] = float(max(0, slack * (mu_k + sum_m c_m G_mk + cb_k[code])))  */
int anqi_lrq_materialize(anqi_ctx* ctx, const float* coeff, const uint8_t* codes_kmajor,
                         uint64_t n, uint32_t nk, uint32_t rankm, const float* mu,
                         const float* basis, const float* codebook, uint32_t code_size,
                         uint32_t k_one_based, double slack, float* thr);
/* The same for nks ranks ks[0..nks) (host, one-based) in one pass over the
 * rows: thr is device nks x n, table t for rank ks[t] (e.g. the batcher's
 * k = 10, 50, 100 tables).  materialize_k once per rank in the reference
 * (src/rknn_query.cpp:722-753). */
int anqi_lrq_materialize_multi(anqi_ctx* ctx, const float* coeff, const uint8_t* codes_kmajor,
                               uint64_t n, uint32_t nk, uint32_t rankm, const float* mu,
                               const float* basis, const float* codebook, uint32_t code_size,
                               const uint32_t* ks, uint32_t nks, double slack, float* thr);
/* The same with the metric: ANQI_METRIC_IP keeps the sign,
 * thr[i] = float(slack * (mu_k + sum_m c_m G_mk + cb_k[code])), as
 * materialize_k does for "ip" (src/rknn_query.cpp:722-753, :751). */
int anqi_lrq_materialize_metric(anqi_ctx* ctx, const float* coeff, const uint8_t* codes_kmajor,
                                uint64_t n, uint32_t nk, uint32_t rankm, const float* mu,
                                const float* basis, const float* codebook, uint32_t code_size,
                                const uint32_t* ks, uint32_t nks, double slack, int metric, float* thr);
/* Arithmetic of the verifier's encode and decode on this context: 0
 * (default) FP32 float-float sums whose rounding to float is certified
 * against the binary64 evaluation's error bound, soft binary64 only for the
 * elements it cannot certify; 1 soft binary64 for every element.  Both give
 * the reference's bits (src/rknn_query.cpp:679-753); form 1 exists to check
 * that.  anqi_ctx_lrq_fallbacks reports how many elements the last encode
 * and the last decode recomputed in soft binary64 (form 0). */
int anqi_ctx_set_lrq_form(anqi_ctx* ctx, int form);
int anqi_ctx_lrq_fallbacks(anqi_ctx* ctx, uint64_t* encode, uint64_t* decode);

/* ---- relaxed verifier fit (tools/train_final_m8_u8.py:81-156 objective) --
 * Radius reconstruction with the codeword choice relaxed to a softmax over
 * the code_size codewords of each rank (temperature tau):
 *   loss = mean_ik (r_ik - mu_k - sum_m c_im G_mk - sum_j w_ikj cb_kj)^2,
 *   c_i = (r_i - mu) G^T,  w_ik = softmax_j(-(e_ik - cb_kj)^2 / tau).
 * mode 1 relaxes the floor decode itself (sorted codewords):
 *   q = cb_0 + sum_{j>=1} logistic((e - cb_j)/tau) (cb_j - cb_{j-1}).
 * All pointers device except *loss (host, may be NULL).  grad_basis
 * (rankm x nk) and grad_codebook (nk x code_size) receive grad_scale times
 * the gradient of the mean loss (either may be NULL).  Hand-written backward;
 * the exact floor decode (anqi_lrq_encode/materialize) rules on the result. */
int anqi_lrq_relax(anqi_ctx* ctx, const float* cutoffs, uint64_t n, uint32_t nk, uint32_t rankm,
                   const float* mu, const float* basis, const float* codebook, uint32_t code_size,
                   float tau, int mode, float grad_scale, double* loss, float* grad_basis,
                   float* grad_codebook);

/* AKNN keys (anqi_aknn_build, raw d2) -> the cutoffs the .rec holds,
 * cut = max(0, keys) (tools/build_aknn_source.cpp:119-137), and the lift's
 * horizon[i] = cut[i][K-1] (src/rknn_query.cpp:2203-2236).  device keys and
 * cut (n*K, may alias), horizon (n, may be NULL). */
int anqi_aknn_cutoffs(anqi_ctx* ctx, const float* keys, uint64_t n, uint32_t K, float* cut, float* horizon);
/* The same for the metric: ANQI_METRIC_IP gives cut = keys - 1 (the "ip"
 * .rec key, tools/build_aknn_source.cpp:125-126). */
int anqi_aknn_cutoffs_metric(anqi_ctx* ctx, const float* keys, uint64_t n, uint32_t K, int metric, float* cut,
                             float* horizon);

/* ---- lift (src/rknn_query.cpp:2203-2236), host FP64 ----------------------- */
/* host base (n*d), rk_horizon (n); host lifted (n*(d+2)). */
int anqi_lift_host(const float* base, uint64_t n, uint32_t d, const float* rk_horizon,
                   float* lifted, double* mstar2, double* scale);
/* The same lift on the device (rknn_query.cpp:2203-2236): device base (n*d),
 * rk_horizon (n), lifted (n*(d+2)).  Per-row binary64 work runs in integer
 * soft binary64 (no FP64 instructions); the row-order sum of squared norms is
 * taken on the host.  Same bits as anqi_lift_host. */
int anqi_lift(anqi_ctx* ctx, const float* base, uint64_t n, uint32_t d, const float* rk_horizon, float* lifted,
              double* mstar2, double* scale);
/* Both lifts with the metric: ANQI_METRIC_IP sets the lift coordinate to
 * float(r * s^2) instead of float(0.5 (r s^2 - |s o|^2)) (src/rknn_query.cpp:
 * 2227); everything else is shared.  The first two are the L2 forms above. */
int anqi_lift_host_metric(const float* base, uint64_t n, uint32_t d, const float* rk_horizon, int metric,
                          float* lifted, double* mstar2, double* scale);
int anqi_lift_metric(anqi_ctx* ctx, const float* base, uint64_t n, uint32_t d, const float* rk_horizon,
                     int metric, float* lifted, double* mstar2, double* scale);

/* ---- graph construction pieces (include/index.h:1071-1295) ---------------- */
int anqi_warm_start(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d,
                    const uint32_t* ids, uint32_t width, uint32_t cap, uint32_t* hidx,
                    float* hval);
/* LocalJoin, snapshot schedule; upd (host, iters entries) receives updates
 * per pass; returns the pass count through *passes.  form 1: the sample's
 * pair distances as a bf16 MFMA tile with FP32 re-scoring near the roots
 * (d <= 144, sample <= 32); same result as form 0. */
int anqi_nndescent(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d,
                   uint32_t cap, uint32_t* hidx, float* hval, int iters, int sample,
                   int64_t* upd, int* passes, int form);
/* Reverse lists of a rows x width id table (rowcnt NULL: sentinel-filtered). */
int anqi_reverse_csr(anqi_ctx* ctx, const uint32_t* ids, uint32_t rows, uint32_t width,
                     const uint32_t* rowcnt, uint32_t n, uint32_t* roff, uint32_t* rids);
int anqi_gather_topk(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d,
                     const uint32_t* hidx, const float* hval, uint32_t cap,
                     const uint32_t* roff, const uint32_t* rids, uint32_t explore,
                     uint32_t pool_cap, uint32_t K, uint32_t* ids, float* keys,
                     uint32_t* cnt, int form);
int anqi_gather_prune(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d,
                      const uint32_t* hidx, const float* hval, uint32_t cap,
                      const uint32_t* roff, const uint32_t* rids, uint32_t explore,
                      uint32_t pool_cap, uint32_t M, float alpha, uint32_t* O,
                      uint32_t* ocnt, int form);   /* O slots past ocnt[i]: UINT32_MAX */
int anqi_commit(anqi_ctx* ctx, const float* x, const float* nrm, uint32_t n, uint32_t d,
                const uint32_t* O, const uint32_t* ocnt, uint32_t M, float alpha,
                uint32_t* links, uint32_t* deg);

/* ---- search tree (include/rptrees.h:108-184) ------------------------------- */
typedef struct {
    uint32_t nodes, root, d;
    uint32_t* left;    /* device; UINT32_MAX for leaves */
    uint32_t* right;
    float* offset;
    float* hyper;      /* nodes * d */
    uint32_t* entry;   /* leaf: first member (indices[0]) */
    uint64_t* leaf_off;  /* host: nodes+1 offsets into leaf_ids (export) */
    uint32_t* leaf_ids;  /* host: members of every leaf in stored order */
} anqi_tree;
int anqi_rp_tree(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, uint32_t leaf,
                 uint32_t max_depth, uint64_t seed, anqi_tree* out);
/* Upload a tree given in host arrays (e.g. read from a reference graph file). */
int anqi_tree_upload(anqi_ctx* ctx, uint32_t nodes, uint32_t root, uint32_t d,
                     const uint32_t* left, const uint32_t* right, const float* offset,
                     const float* hyper, const uint32_t* entry, anqi_tree* out);
int anqi_tree_free(anqi_ctx* ctx, anqi_tree* t);

/* An edge policy of the graph's prunes (include/index.h:283-829) as
 * configure_edge_residual_cover (:305-438) leaves it: every prune of the build
 * (the gathered pools, commit_edges and the Vamana passes) follows
 * prune_pool_to_ids's branch for these flags.  All flags zero: the plain
 * alpha-RNG prune. */
typedef struct {
    int resid_cover;    /* "residual_cover" */
    int ball;           /* "ball_*", "bubble_*", "lsg_bubble_hybrid" */
    int ball_pure;      /* "ball_rng", "bubble_rng" */
    int occlude_all;    /* ANQI_EDGE_BALL_OCCLUSION other than "any" */
    int lsg;            /* "lsg_rng", "lsg_bubble_hybrid" */
    double lsg_alpha;   /* ANQI_LSG_ALPHA, 1 */
    const float* mu;    /* device n: ANQI_LSG_MU_FILE */
    const float* resid; /* device n*nk: the residual table */
    uint32_t nk;        /* ANQI_EDGE_NK, 100 */
    int32_t ks[16];     /* ANQI_EDGE_KS, 1-based ranks <= nk */
    uint32_t nks;       /* at most 16 */
    uint32_t budget;    /* min(ANQI_EDGE_RESID_BUDGET, M), 16 */
    double lambda;      /* ANQI_EDGE_RESID_LAMBDA, 1 */
    double scale2;      /* the lift scale squared (_anqi_meta.txt), 1 when absent */
    int bubble_auto;    /* "bubble_auto", "bubble_auto_full" (with ball set) */
    int bubble_full;    /* "bubble_auto_full": every pool entry an alternative and a target */
    uint32_t auto_candidates;  /* ANQI_EDGE_AUTO_CANDIDATES, M */
    uint32_t auto_max_swaps;   /* ANQI_EDGE_AUTO_MAX_SWAPS, M */
} anqi_edge_policy;

typedef struct {
    uint32_t M;        /* graph degree 48 */
    uint32_t nn_k;     /* LIFTNNK 50 */
    uint32_t leaf;     /* search tree leaf 100 */
    uint32_t explore;  /* 100 */
    uint32_t ef_alpha; /* 8 */
    int iters;         /* GRAPHNND 6 */
    int sample;        /* NNSAMPLE 20 */
    float alpha;       /* 1.2 */
    uint64_t seed;     /* 10 */
    int form;          /* 0 plain (per-chunk gather), 1 chosen (per-expansion
                          batched gather); both give the same graph */
    /* Vamana refinement after the commit (include/index.h:1262-1282; zero:
     * off).  1: ANQI_VAMANA_STYLE=global -- each pass collects every row's
     * beam from its own tree entry over the current graph, adds its heap
     * slots, prunes and commits (:1267-1279); 2: "batch" --
     * vamana_batch_refine (:932-1063): rows in a shuffled order, in growing
     * batches from one entry, each batch's rows and the rows they point to
     * re-pruned before the next batch. */
    int vamana;
    uint32_t vamana_L;            /* ANQI_VAMANA_L, 128; beam width <= 1024 */
    int vamana_passes;            /* ANQI_VAMANA_PASS, 2 */
    double vamana_batch_frac;     /* ANQI_VAMANA_BATCH_FRAC, 0.02 */
    uint64_t vamana_max_batch;    /* ANQI_VAMANA_MAX_BATCH, 1000000 */
    double vamana_batch_base;     /* ANQI_VAMANA_BATCH_BASE, 2 */
    float vamana_early_alpha;     /* ANQI_VAMANA_EARLY_ALPHA, 1.0: the prune alpha before the last pass */
    int vamana_early_alpha_one;   /* ANQI_VAMANA_EARLY_ALPHA_ONE, 1 */
    uint32_t vamana_entry;        /* the batch style's entry (choose_vamana_entry: 0, an id
                                     mod n, or anqi_vamana_medoid's row) */
    const anqi_edge_policy* edge; /* NULL: the plain prune */
} anqi_graph_params;
typedef struct {
    double tree_s, warm_s, nnd_s, gather_s, commit_s, total_s;
    int nnd_passes;
    int64_t nnd_upd[16];
    uint32_t overflow_rows;
    uint64_t gather_d2, gather_adj;  /* gather work: distances computed, adjacency ids read */
    double vamana_s;                 /* Vamana passes */
} anqi_graph_stats;
/* choose_vamana_entry "medoid" (include/index.h:190-222): the sampled row
 * (rows (t * step) % n, step = max(1, n / sample), t < sample, sample
 * clamped to [1, n]) nearest, in binary64, to the sample's binary64 mean.
 * x: device rows (the lifted rows for the graph). */
int anqi_vamana_medoid(anqi_ctx* ctx, const float* x, uint32_t n, uint32_t d, uint64_t sample, uint32_t* entry);
/* device lifted (n*dl), warm_ids (n*width); device links (n*M), deg (n);
 * link slots past deg[i] hold UINT32_MAX. */
int anqi_graph_build(anqi_ctx* ctx, const float* lifted, uint32_t n, uint32_t dl,
                     const uint32_t* warm_ids, uint32_t width, const anqi_graph_params* p,
                     uint32_t* links, uint32_t* deg, anqi_tree* tree, anqi_graph_stats* st);

/* ---- RkNN query (index.h:1535-1652 + rknn_query.cpp:2692-2824) ----------- */
typedef struct {
    const float* base;      /* device n*d */
    const float* bnorm;     /* device n */
    const float* lifted;    /* device n*dl */
    const uint32_t* deg;    /* device n */
    const uint32_t* links;  /* device n*M */
    uint32_t n, d, dl, M;
    const anqi_tree* tree;
    double mstar2, scale;
    /* optional split lifted layout from anqi_lift_split (both NULL: the
     * persistent search reads the lifted rows); used when d % 16 == 0 */
    const float* lmain;     /* device n*d */
    const float* ltail;     /* device n */
    /* optional, with lmain/ltail: sentinel adjacency from anqi_pack_links
     * (NULL: the search reads links and deg) */
    const uint32_t* plinks; /* device n*M */
    /* optional, with lmain/ltail: byte rows from anqi_byte_rows when it
     * reports them exact (NULL: the search reads float rows) */
    const uint8_t* brows;   /* device n*d */
    /* ANQI_METRIC_L2 (0) or ANQI_METRIC_IP: the inner-product search takes
     * each query's R2 = |s q|^2 + 1 + M*^2 and the post-filter key -<o, q>
     * (src/rknn_query.cpp:2555-2561, 2719-2730); the lifted rows must come
     * from the lift of the same metric */
    int metric;
} anqi_index;
/* Split lifted layout for the persistent search: lmain = the first d
 * coordinates of each lifted row (rows of d floats, so a row covers d/32
 * whole 128-byte lines when d % 16 == 0), ltail = each row's lift-coordinate
 * term of the lifted distance to a query's (1, 0), reduced as the whole-row
 * distance reduces it.  D = main term + ltail[i] has the bits of the
 * whole-row D.  No counterpart in the reference (a layout of
 * rknn_query.cpp:2203-2236's lifted rows).  d % 16 == 0. */
int anqi_lift_split(anqi_ctx* ctx, const float* lifted, uint32_t n, uint32_t d, float* lmain, float* ltail);
/* Byte rows for the persistent search over the split layout, for a base
 * whose values are all the floats 0..255 (SIFT-like data): brows holds each
 * row's d values as bytes, ordered for the search's 16-lane groups.  The
 * search rebuilds the base value float(b) and, in FP32, the lifted
 * coordinate -- the lift's float(double(b) * scale), rknn_query.cpp:
 * 2203-2236 -- from them, reading d bytes per distance instead of 4d.
 * *exact = 1 iff every element's base value and lmain coordinate are
 * reproduced bit for bit; otherwise brows must not be used.
 * No counterpart in the reference (a layout of its base and lifted rows).
 * d % 16 == 0; scale: the lift's scale (anqi_lift). */
int anqi_byte_rows(anqi_ctx* ctx, const float* base, const float* lmain, uint32_t n, uint32_t d, double scale,
                   uint8_t* brows, int* exact);
/* Sentinel adjacency for the persistent search over the split layout:
 * plinks[i M + j] = links[i M + j] for j < deg[i], else UINT32_MAX, whatever
 * links holds past the degree, so an expansion reads its row and no degree.
 * No counterpart in the reference (a layout of include/index.h:65-69's
 * links for index.h:1535-1652's expansions).  M <= 64. */
int anqi_pack_links(anqi_ctx* ctx, const uint32_t* links, const uint32_t* deg, uint32_t n, uint32_t M,
                    uint32_t* plinks);
/* queries (device nq*d, base space).  thr: device thresholds for k (n floats)
 * or, with kidx, tables (ntables*n) picked per query.  Lq/kidx optional
 * device per-query arrays.  Outputs (device): cand_ids/accept nq*out_cap,
 * cand_cnt, ndist (graph distances), pred_cnt, nexp (expanded nodes,
 * optional), status (0 ok, 3 truncated).
 * mode 0: one wavefront per query; 1: persistent work-queue kernel. */
int anqi_rknn_search(anqi_ctx* ctx, const anqi_index* ix, const float* queries, uint32_t nq,
                     uint32_t L, const uint32_t* Lq, uint32_t Lmax, const float* thr,
                     const uint32_t* kidx, uint32_t out_cap, uint32_t* cand_ids,
                     uint8_t* accept, uint32_t* cand_cnt, uint32_t* ndist,
                     uint32_t* pred_cnt, uint32_t* nexp, uint32_t* status, int mode);
/* Same, with explicit per-query entry points (device, nq) in place of the
 * search-tree descent (the reference's ANQI_QUERY_SINGLE_ENTRY analogue,
 * index.h:1597-1598); used to isolate graph quality from entry quality. */
int anqi_rknn_search_entries(anqi_ctx* ctx, const anqi_index* ix, const float* queries, uint32_t nq,
                             uint32_t L, const uint32_t* Lq, uint32_t Lmax, const float* thr,
                             const uint32_t* kidx, const uint32_t* entries, uint32_t out_cap,
                             uint32_t* cand_ids, uint8_t* accept, uint32_t* cand_cnt,
                             uint32_t* ndist, uint32_t* pred_cnt, uint32_t* nexp,
                             uint32_t* status, int mode);

/* Persistent form of anqi_rknn_search (mode 1) that allocates nothing: its
 * counters, per-query batch arrays, redo list and global-tier rows live in a
 * 64 MiB work area the library holds on each device for the life of the
 * process, so callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation) can
 * search.  No counterpart in the reference (its query loop allocates per
 * thread, src/rknn_query.cpp:2692-2824).  Same arguments, outputs and
 * statuses as anqi_rknn_search with mode 1 and the context's search tuning.
 * Limits: nq <= 65536 per call and n <= about 16.0M rows (one global-tier
 * bitmap + FIFO row pair); larger inputs return ANQI_E_USAGE before anything
 * runs.  Calls that overlap on one device share the work area and race. */
int anqi_rknn_search_noalloc(anqi_ctx* ctx, const anqi_index* ix, const float* queries, uint32_t nq,
                             uint32_t L, const uint32_t* Lq, uint32_t Lmax, const float* thr,
                             const uint32_t* kidx, uint32_t out_cap, uint32_t* cand_ids, uint8_t* accept,
                             uint32_t* cand_cnt, uint32_t* ndist, uint32_t* pred_cnt, uint32_t* nexp,
                             uint32_t* status);

/* The candidates of a search's dense outputs as CSR: off (device, nq+1) =
 * exclusive scan of min(cand_cnt, out_cap); ids / acc_out (device,
 * *total entries) = each query's candidate ids and accept bytes in
 * discovery order.  *total is returned first; with ids == NULL only off and
 * *total are computed (size the buffers, then call again).  The reference
 * keeps per-query vectors (src/rknn_query.cpp:2705-2776); this is the form a
 * host reads them in. */
int anqi_compact_candidates(anqi_ctx* ctx, uint32_t nq, uint32_t out_cap, const uint32_t* cand_ids,
                            const uint8_t* accept, const uint32_t* cand_cnt, uint32_t* off, uint32_t* ids,
                            uint8_t* acc_out, uint64_t ids_cap, uint64_t* total);

/* Search-tree entry point of each query (device nq): the descent the search
 * runs (rptrees.h:108-123, ties go left), for anqi_rknn_search_entries. */
int anqi_search_entries(anqi_ctx* ctx, const anqi_index* ix, const float* queries, uint32_t nq, uint32_t* entries);

/* ---- captured query batches (configuration 5: mixed k and L) -------------
 * No counterpart in the reference (its query driver loops over queries on
 * CPU threads, src/rknn_query.cpp:2692-2824): this is the batch form of the
 * same range search + post-filter, for serving.
 * A batcher owns device buffers for up to max_batch queries and an
 * instantiated hipGraph {counter reset, persistent search, answer scan,
 * answer compaction}; the batch size is read on the device, so the same graph
 * serves every batch.  thr_tables: device, ntables x n thresholds (one table
 * per k, from anqi_lrq_materialize); a query's kidx selects its table and its
 * L (1..Lmax) its beam width.  anqi_batcher_run copies host queries in, runs
 * the graph, reruns (uncaptured) the rare queries that outgrow the tuned work
 * areas, and returns the accepted ids of every query as host CSR
 * (h_off[nq+1], h_ids[*total], candidate order). The index and tables must
 * outlive the batcher; one thread per batcher. */
typedef struct anqi_batcher anqi_batcher;
int anqi_batcher_create(anqi_ctx* ctx, const anqi_index* ix, const float* thr_tables, uint32_t ntables,
                        uint32_t max_batch, uint32_t Lmax, uint32_t out_cap,
                        const anqi_search_tuning* tune, anqi_batcher** out);
int anqi_batcher_run(anqi_batcher* b, const float* h_queries, const uint32_t* h_L, const uint32_t* h_kidx,
                     uint32_t nq, uint32_t* h_off, uint32_t* h_ids, uint64_t ids_cap, uint64_t* total);
void anqi_batcher_destroy(anqi_batcher* b);
/* The same counts for the batcher's last run (the global tier also takes the
 * queries whose candidates exceed out_cap; their answers are complete). */
int anqi_batcher_redo_counts(const anqi_batcher* b, uint32_t* retry, uint32_t* global);
/* Times the persistent search on a sample batch (device queries, Lq, kidx)
 * for visited-set sizes {1536, 2048, 3072, 3584, 4096, 8192}, LDS FIFO sizes {128, 512} and
 * 1/2, 3/4, 4/4 of the resident waves per CU; queries that overflow are
 * charged as reruns.  Writes the fastest tuning and its time. */
int anqi_search_autotune(anqi_ctx* ctx, const anqi_index* ix, const float* thr_tables,
                         const float* queries, uint32_t nq, const uint32_t* Lq,
                         const uint32_t* kidx, uint32_t Lmax, anqi_search_tuning* out,
                         double* best_ms);

/* Library build description (compiler, arch). */
const char* anqi_build_info(void);

#ifdef __cplusplus
}
#endif
#endif
