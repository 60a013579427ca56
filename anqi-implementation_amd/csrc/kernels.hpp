// kernels.hpp -- parameter blocks and host launchers shared between the
// kernel translation units and the C-ABI layer (abi.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

// ---- range search + postfilter (k_search.hip) ------------------------------
struct SearchParams {
    const float* base;      // n x d
    const float* bnorm;     // n
    const float* lifted;    // n x dl
    const uint32_t* deg;    // n
    const uint32_t* links;  // n x M
    uint32_t n, d, dl, M;
    const uint32_t* t_left;
    const uint32_t* t_right;
    const float* t_off;
    const float* t_hyper;   // nodes x dl
    const uint32_t* t_entry;
    uint32_t t_root;
    const uint32_t* entries;   // optional per-query entry points (tree bypassed)
    uint64_t scale_bits;    // binary64 lift scale
    float R2f;              // largest float <= M*^2 + 1
    const float* qr2;       // inner product: per-query R2f (k_query_r2); null: R2f for every query
    int ip;                 // inner-product metric: post-filter key -<o, q>
    const float* queries;   // nq x d (base space)
    uint32_t nq;
    const uint32_t* qlist;  // optional query index list
    uint32_t L, Lmax;
    const uint32_t* Lq;     // optional per-query L
    const float* thr;       // thresholds (one table, or tables x n)
    const uint32_t* kidx;   // optional per-query table index
    uint32_t hmask;         // visited hash size - 1
    uint32_t rcap;          // LDS FIFO capacity
    uint32_t out_cap;       // per-query output capacity
    uint32_t* cand_ids;     // nq x out_cap
    uint8_t* accept;        // nq x out_cap
    uint32_t* cand_cnt;     // nq
    uint32_t* ndist;        // nq (graph distance evaluations)
    uint32_t* pred_cnt;     // nq
    uint32_t* nexp;         // nq: expanded nodes (optional)
    uint32_t* status;       // nq: 0 ok, 1 hash full, 2 FIFO full, 3 output truncated
    uint32_t* work_counter; // persistent form
    uint32_t* ovf_count;    // optional: number of queries that overflowed (status 1 or 2)
    const uint32_t* nq_dev; // optional: batch size read on the device (captured batches)
    uint32_t rbeam;         // set by launch_range_search: register beam (persistent form)
    // split lifted layout (persistent form, d % 16 == 0; null: lifted rows):
    const float* lmain;     // n x d: the lifted rows' first d coordinates
    const float* ltail;     // n: reduced term of each row's lift coordinates against a query's (1, 0)
    // sentinel adjacency (with lmain only; null: links + deg): n x M ids,
    // U32_NONE past the degree, so an expansion reads no degree
    const uint32_t* plinks;
    // byte rows (with lmain only; null: float rows): n x d bytes, lane-ordered
    // (k_search.hip brow_fetch), certified by launch_byte_rows
    const uint8_t* brows;
    // global-memory tier (k_range_search_global): block b runs query
    // qlist[b] with a visited bitmap of n bits at gvis + b * gstride words and
    // a candidate FIFO of n entries at gfifo + b * n, so no query outgrows it
    uint32_t* gvis;
    uint32_t* gfifo;
    uint32_t gstride;       // bitmap words per block (a multiple of 4)
};
// Global tier: one block per listed query (P.qlist[0..P.nq)), P.gvis/gfifo
// sized for P.nq blocks.  Statuses end 0 or 3 (more than out_cap candidates);
// each block's FIFO row then holds the query's whole candidate list in
// discovery order, bit 31 set on the accepted ones.
hipError_t launch_range_search_global(const SearchParams& P, hipStream_t s);
// Indices q < nq whose status[q] is in the set `mask` (bit v: status v) into
// list, their count into *count (one block; list order unspecified).
hipError_t launch_status_list(const uint32_t* status, uint32_t nq, uint32_t mask, uint32_t* list, uint32_t* count,
                              hipStream_t s);
// Lift of the base rows (k_lift.hip): pass 1 per-row squared norms (binary64
// bits), pass 2 rows h_0..h_d and the maximum lifted norm (atomic maximum on
// the bits of *maxbits, which starts at 0), pass 3 the padding coordinate.
hipError_t launch_lift_norm2(const float* x, uint64_t n, uint32_t d, uint64_t* s_out, hipStream_t s);
hipError_t launch_lift_rows(const float* x, uint64_t n, uint32_t d, const float* rk, uint64_t scale, uint64_t sc2,
                            int ip, float* lifted, uint64_t* maxbits, hipStream_t s);
hipError_t launch_lift_pad(float* lifted, uint64_t n, uint32_t d, uint64_t m2, hipStream_t s);
// cut = max(0, keys) (n x K), horizon (optional, n) = cut[:, K-1].
hipError_t launch_aknn_cutoffs(const float* keys, uint64_t n, uint32_t K, int ipm, float* cut, float* horizon,
                               hipStream_t s);
// Byte rows of a base whose values are all floats 0..255: brows as
// brow_fetch reads them; *bad (zeroed by the caller) counts elements whose
// base value or lifted coordinate (lmain) the bytes do not reproduce.
hipError_t launch_byte_rows(const float* base, const float* lmain, uint32_t n, uint32_t d, double scale,
                            uint8_t* brows, uint32_t* bad, hipStream_t s);
// Sentinel adjacency for the split layout: plinks[i M + j] = links[i M + j]
// for j < deg[i], else U32_NONE.
hipError_t launch_pack_links(const uint32_t* links, const uint32_t* deg, uint32_t n, uint32_t M, uint32_t* plinks,
                             hipStream_t s);
// Split lifted layout of an index: lmain (n x d), ltail (n).
hipError_t launch_lift_split(const float* lifted, uint32_t n, uint32_t d, float* lmain, float* ltail,
                             hipStream_t s);
size_t search_lds_bytes(const SearchParams& P);
int search_persistent_occupancy(const SearchParams& P);
hipError_t launch_range_search(const SearchParams& P, bool persistent, uint32_t grid, hipStream_t s);
// Search-tree entries of queries [0, nq) (nq from P.nq_dev when set) into out;
// key (optional): each query's lifted distance to its entry.
hipError_t launch_tree_entry_batch(const SearchParams& P, uint32_t max_nq, uint32_t* out, float* key,
                                   hipStream_t s);
// Work-queue order (qlist) of queries [0, nq) by ascending key (one workgroup).
hipError_t launch_entry_order(const SearchParams& P, const float* key, uint32_t* qlist, hipStream_t s);
// Inner-product metric: per-query R2f (largest float <= |s q|^2 + 1 + M*^2)
// into r2[0..nq) (nq from *P.nq_dev when set; max_nq threads launched).
hipError_t launch_query_r2(const SearchParams& P, uint32_t max_nq, double mstar2, float* r2, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s);
// Batch result compaction (k_batch.hip): off = exclusive scan of pred_cnt over
// *nq_dev queries (one block), then each query's accepted ids in candidate
// order into ids[off[q] ..].  Capturable (no host reads).
hipError_t launch_batch_scan(const uint32_t* pred_cnt, const uint32_t* nq_dev, uint32_t max_nq, uint32_t* off,
                             hipStream_t s);
hipError_t launch_batch_compact(const uint32_t* cand_ids, const uint8_t* accept, const uint32_t* cand_cnt,
                                const uint32_t* off, const uint32_t* nq_dev, uint32_t max_nq, uint32_t out_cap,
                                uint32_t* ids, hipStream_t s);
// Candidate CSR (k_batch.hip): off = exclusive scan of min(cnt, out_cap),
// then each query's candidate ids and accept bytes.
hipError_t launch_cand_scan(const uint32_t* cnt, uint32_t nq, uint32_t out_cap, uint32_t* off, hipStream_t s);
hipError_t launch_cand_copy(const uint32_t* cand, const uint8_t* acc, const uint32_t* off, uint32_t nq,
                            uint32_t out_cap, uint32_t* ids, uint8_t* acc_out, hipStream_t s);

// ---- verifier (k_lrq.hip) ---------------------------------------------------
hipError_t launch_lrq_encode(const float* cut, uint64_t n, uint32_t nk, uint32_t rankm,
                             const float* mu, const float* basis, const float* cb,
                             uint32_t code_size, float* coeff, uint8_t* codes,
                             uint8_t* codes_kmajor, hipStream_t s);
hipError_t launch_lrq_materialize(const float* coeff, const uint8_t* codes_kmajor, uint64_t n,
                                  uint32_t nk, uint32_t rankm, const float* mu,
                                  const float* basis, const float* cb, uint32_t code_size,
                                  uint32_t k_one_based, uint64_t slack_bits, int clamp, float* thr,
                                  hipStream_t s);
hipError_t launch_transpose_codes(const uint8_t* codes, uint64_t n, uint32_t nk,
                                  uint8_t* codes_kmajor, hipStream_t s);
// FP32 float-float forms with certified rounding and soft-binary64 fallback
// for the rare uncertified element (same bits as the soft forms above);
// fallbacks (optional, device) counts the recomputed elements.
#define LRQ_MAX_TABLES 8
bool lrq_fast_supported(uint32_t nk, uint32_t rankm, uint32_t code_size);
hipError_t launch_lrq_encode_fast(const float* cut, uint64_t n, uint32_t nk, uint32_t rankm, const float* mu,
                                  const float* basis, const float* cb, float* coeff, uint8_t* codes,
                                  uint8_t* codes_kmajor, unsigned long long* fallbacks, hipStream_t s);
// Thresholds of nks ranks (host list, one-based) into thr (nks x n), slack 1.
hipError_t launch_lrq_decode(const float* coeff, const uint8_t* codes_kmajor, uint64_t n, uint32_t nk,
                             uint32_t rankm, const float* mu, const float* basis, const float* cb,
                             uint32_t code_size, const uint32_t* k_one_based, uint32_t nks, int clamp,
                             float* thr, unsigned long long* fallbacks, hipStream_t s);

// ---- distances (k_dist.hip) ---------------------------------------------------
hipError_t launch_norms(const float* x, uint64_t n, uint32_t d, float* out, hipStream_t s);
hipError_t launch_pair_d2(const float* x, const float* nrm, uint32_t d, const uint32_t* a,
                          const uint32_t* b, uint64_t m, float* out, hipStream_t s);

// ---- graph construction (k_graph.hip) ------------------------------------------
hipError_t launch_warm_start(const float* x, const float* nrm, uint32_t n, uint32_t d,
                             const uint32_t* ids, uint32_t width, uint32_t cap,
                             uint32_t* hidx, float* hval, hipStream_t s);
hipError_t launch_nnd_sample(const uint32_t* hidx, uint32_t n, uint32_t cap, uint32_t sample,
                             uint32_t* samp, uint32_t* scnt, hipStream_t s);
hipError_t launch_nnd_propose(const float* x, const float* nrm, uint32_t n, uint32_t d,
                              const uint32_t* samp, const uint32_t* scnt, uint32_t sample,
                              const float* hval, uint32_t cap, uint64_t* keys, uint32_t* vals,
                              unsigned long long* count, uint64_t max_props, hipStream_t s);
// Chosen form (k_graph.hip): bf16 MFMA gram of the sample, FP32 re-score of
// the pairs within the bound marg (k_exact_margin of the rows); same proposals.
bool nnd_propose_mfma_supported(uint32_t d, uint32_t sample);
hipError_t launch_nnd_propose_mfma(const float* x, const float* nrm, const float* marg, uint32_t n, uint32_t d,
                                   const uint32_t* samp, const uint32_t* scnt, uint32_t sample, const float* hval,
                                   uint32_t cap, uint64_t* keys, uint32_t* vals, unsigned long long* count,
                                   uint64_t max_props, hipStream_t s);
hipError_t launch_nnd_apply(const uint64_t* keys, const uint32_t* vals, uint64_t m,
                            const uint32_t* row_off, uint32_t n, uint32_t* hidx, float* hval,
                            uint32_t cap, unsigned long long* upd, hipStream_t s);
hipError_t launch_row_bounds(const uint64_t* keys, uint64_t m, uint32_t n, uint32_t* row_off,
                             hipStream_t s);
hipError_t launch_count_refs(const uint32_t* ids, uint64_t m, uint32_t n, uint32_t* cnt,
                             hipStream_t s);
hipError_t launch_scatter_rev(const uint32_t* ids, uint32_t rows, uint32_t width,
                              const uint32_t* rowcnt, uint32_t n, uint32_t* cursor,
                              uint32_t* rids, hipStream_t s);

// An edge policy of the graph prunes (include/index.h:283-829), configured
// (configure_edge_residual_cover, :305-438); on == 0: the plain prune.
struct EdgePolicyDev {
    int on;                     // 0: the plain alpha-RNG prune
    int resid_cover, ball, ball_pure, occlude_all, lsg;
    double lsg_alpha, lambda, scale2;
    const float* mu;            // n lsg weights
    const float* resid;         // n x nk residuals
    uint32_t nk, nks, budget;
    int32_t ks[16];             // 1-based ranks
    int bubble_auto, bubble_full;              // optimize_verification_constrained_bubble
    uint32_t auto_candidates, auto_max_swaps;
};

struct GatherParams {
    const float* x;
    const float* nrm;
    uint32_t n, d;
    const uint32_t* hidx;   // n x cap
    const float* hval;
    uint32_t cap;
    const uint32_t* roff;   // n+1
    const uint32_t* rids;
    uint32_t explore;       // queue capacity
    uint32_t pool_cap;      // stop once the pool holds this many
    uint32_t K;             // top-K output (AKNN mode)
    uint32_t M;             // prune degree (graph mode)
    float alpha;
    uint32_t hmask;         // visited hash size - 1
    uint32_t pmax;          // pool capacity in the work area (graph mode)
    const uint32_t* rows;   // optional row list
    uint32_t nrows;
    uint32_t* out_ids;      // AKNN: n x K ; graph: n x M
    float* out_keys;        // AKNN: n x K
    uint32_t* out_cnt;      // n
    uint32_t* status;       // n: 0 ok, 1 hash full, 2 pool full
    uint32_t* gscratch;     // optional global work area (per block)
    size_t gscratch_words;  // words per block
    uint32_t batched;       // 1: one expansion's first visits processed together
    uint32_t h16;           // 16-bit LDS visited set of H16_SLOTS entries (hmask unused): 1 n <= 2^20,
                            // 2 n <= 2^24 (with a 32-bit overflow table); 0: 32-bit entries
    unsigned long long* work;  // optional: [0] += distances computed, [1] += adjacency ids read (finished rows)
    int ipm;                // AKNN "ip" metric: distance 1 - <x_i, x_c> (ipdist_seq), one lane per pair
    // Vamana collect (graph mode): vmode 1 -- the beam starts at ventry[i]
    // (or ventry0) and expands the fixed-degree graph vlinks/vdeg (row width
    // vM) instead of the heaps; the row's heap slots join the pool after the
    // beam; vmode 2 -- its current edges as well
    int vmode;
    const uint32_t* vlinks;
    const uint32_t* vdeg;
    uint32_t vM;
    const uint32_t* ventry;
    uint32_t ventry0;
    EdgePolicyDev pol;      // graph mode: the prune's edge policy (work area: global only)
};
// rk: register-list form of the batched gather (see k_graph.hip:gather_rk).
bool gather_register_lists(const GatherParams& P, bool graph_mode);
size_t gather_lds_bytes(const GatherParams& P, bool graph_mode, bool rk);
hipError_t launch_gather(const GatherParams& P, bool graph_mode, bool global_ws, hipStream_t s);

struct PruneParams {
    const float* x;
    const float* nrm;
    uint32_t n, d;
    const uint32_t* O;      // n x M (first ocnt[i] valid)
    const uint32_t* ocnt;
    const uint32_t* roff;   // reverse CSR of O
    const uint32_t* rids;
    uint32_t M;
    float alpha;
    uint32_t pmax;          // pool capacity
    uint32_t* links;        // n x M
    uint32_t* deg;
    uint32_t* status;
    const uint32_t* rows;
    uint32_t nrows;
    uint32_t* gscratch;
    size_t gscratch_words;
    const uint32_t* rmap;   // optional: rids hold indices into rmap (the ids are rmap[r])
    EdgePolicyDev pol;      // the prune's edge policy (work area: global only)
};
size_t commit_lds_bytes(const PruneParams& P);
hipError_t launch_commit(const PruneParams& P, bool global_ws, hipStream_t s);

// ---- Vamana refinement (k_graph.hip) ---------------------------------------------
// Search-tree entry of each row of x (n x dl; the rows themselves descend).
hipError_t launch_tree_entry_rows(const float* x, uint32_t n, uint32_t dl, const uint32_t* left,
                                  const uint32_t* right, const float* off, const float* hyper,
                                  const uint32_t* entry, uint32_t root, uint32_t* out, hipStream_t s);
// write_ids for a batch: links/deg rows of rows[0..nb) from O/ocnt (n x M,
// indexed by row id), slots past the degree UINT32_MAX; the batch's rows also
// compacted into Ob/obcnt (nb x M) for the reverse lists.
hipError_t launch_vamana_write(const uint32_t* rows, uint32_t nb, const uint32_t* O, const uint32_t* ocnt,
                               uint32_t M, uint32_t* links, uint32_t* deg, uint32_t* Ob, uint32_t* obcnt,
                               hipStream_t s);
// rows i with roff[i+1] > roff[i] into list (order unspecified), count in *cnt (zeroed here).
hipError_t launch_nonempty_rows(const uint32_t* roff, uint32_t n, uint32_t* list, uint32_t* cnt, hipStream_t s);
// choose_vamana_entry's sampled medoid: mean (binary64, sequential per
// coordinate) of rows (t step) % n, t < sample, then the first sampled row
// of least fused binary64 squared distance.  work: dl doubles + 2 words.
hipError_t launch_vamana_medoid(const float* x, uint32_t n, uint32_t dl, uint64_t sample, double* mean,
                                unsigned long long* best, uint32_t* out, hipStream_t s);

// ---- RP forest + leaf brute force (k_forest.hip) --------------------------------
struct ForestLevel {
    const float* x;
    uint32_t n, d;
    uint32_t* perm;        // trees x n
    uint32_t* perm_out;
    const uint32_t* seg_of; // trees x n: segment id of each position (UINT32_MAX: settled)
    const uint32_t* seg_start;
    const uint32_t* seg_size;
    const uint32_t* seg_pick;  // 2 per segment: picked positions
    float* seg_hyper;       // segments x d
    float* seg_off;
    uint8_t* side;          // trees x n
    uint32_t nseg;
    uint64_t total;         // trees * n
    int angular;            // the "ip" metric's split (rptrees.h:41-52): unit-difference normal, offset 0
};
hipError_t launch_split_planes(const ForestLevel& F, hipStream_t s);
hipError_t launch_split_sides(const ForestLevel& F, uint64_t seed, hipStream_t s);

struct LeafParams {
    const float* x;
    const float* nrm;
    uint32_t n, d;
    const uint32_t* perm;       // the tree's permutation
    const uint32_t* leaf_start;
    const uint32_t* leaf_size;
    uint32_t nleaves;
    uint32_t K;
    uint32_t* ids;              // n x K sorted by (dist, id), UINT32_MAX padded
    float* keys;
    uint32_t* cnt;              // n
    uint32_t max_leaf;
    unsigned long long* nsurv;  // MFMA sweeps: pairs passing the filter (may be null)
    int ipm;                    // "ip" metric: distance 1 - <x, y> (ipdist_seq), FP32 kernel only
};
size_t leaf_lds_bytes(const LeafParams& P);
hipError_t launch_leaf_knn(const LeafParams& P, bool mfma, hipStream_t s);
// MFMA-filtered leaf all-pairs (k_leaf_mfma.hip); same lists as launch_leaf_knn.
// xq: bf16 copy of x, or (fp8) e4m3(x / scale) with scale a power of two;
// marg: per-row filter bound; cbuf (n x capc, capc >= max_leaf) and ccnt (n,
// zeroed) are the candidate lists between the two passes (left zeroed).
bool leaf_mfma_supported(uint32_t d, uint32_t max_leaf, uint32_t K);
hipError_t launch_leaf_mfma(const LeafParams& P, bool fp8, const void* xq, float scale, const float* marg,
                            uint32_t* cbuf, uint32_t* ccnt, uint32_t capc, hipStream_t s);
hipError_t launch_to_fp8(const float* x, const float* nrm, uint32_t n, uint32_t d, float scale, uint8_t* x8,
                         float* nx, float* ex, hipStream_t s);
hipError_t launch_random_fill(const float* x, const float* nrm, uint32_t n, uint32_t d, uint32_t K,
                              uint64_t seed, int ipm, uint32_t* ids, float* keys, uint32_t* cnt,
                              hipStream_t s);

// ---- exact kNN / RkNN membership (k_exact.hip) ------------------------------------
// ipm != 0: inner-product keys -<x, y> (dataset_gt.cpp:225, 292), else d2.
hipError_t launch_exact_topk(const float* x, const float* nrm, uint32_t n, uint32_t d,
                             uint32_t K, uint32_t row0, uint32_t rows, int ipm, uint32_t* ids,
                             float* keys, hipStream_t s);
hipError_t launch_rknn_count(const float* base, const float* bnorm, uint32_t n,
                             const float* q, uint32_t nq, uint32_t d, int ipm, const float* rk2,
                             uint32_t* counts, hipStream_t s);
hipError_t launch_rknn_fill(const float* base, const float* bnorm, uint32_t n,
                            const float* q, uint32_t nq, uint32_t d, int ipm, const float* rk2,
                            const uint32_t* off, uint32_t* ids, float* keys, hipStream_t s);
hipError_t launch_to_bf16(const float* x, uint64_t m, uint16_t* out, hipStream_t s);
// MFMA-filtered exact top-K (k_exact_mfma.hip); same output as launch_exact_topk.
bool exact_mfma_supported(uint32_t d, uint32_t K);
size_t exact_mfma_lds(uint32_t K);
hipError_t launch_exact_mfma(const float* x, const uint16_t* xb, const float* nrm, const float* marg, uint32_t n,
                             uint32_t d, uint32_t K, uint32_t row0, uint32_t rows, int ipm, uint32_t* ids,
                             float* keys, hipStream_t s);
hipError_t launch_exact_norms2(const float* x, const float* nrm, uint32_t n, uint32_t d, float* nx, float* ex,
                               hipStream_t s);
hipError_t launch_exact_margin(const float* nx, const float* ex, uint32_t n, uint32_t d, float Y, float E,
                               float* marg, hipStream_t s);

hipError_t launch_seg_fill(uint32_t* seg_of, const uint32_t* seg_start, const uint32_t* seg_size,
                           uint32_t nseg, hipStream_t s);
hipError_t launch_seg_count_left(const uint8_t* side, const uint32_t* seg_of, uint64_t total,
                                 uint32_t* left, hipStream_t s);
hipError_t launch_seg_redeal(uint8_t* side, const uint32_t* seg_of, const uint8_t* degen, uint64_t total,
                             uint64_t seed, hipStream_t s);
hipError_t launch_left_flags(const uint8_t* side, const uint32_t* seg_of, uint64_t total, uint32_t* flags,
                             hipStream_t s);
hipError_t launch_seg_partition(const uint32_t* perm, uint32_t* perm_out, const uint8_t* side,
                                const uint32_t* seg_of, const uint32_t* seg_start, const uint32_t* left,
                                const uint32_t* lscan, uint64_t total, hipStream_t s);

// ---- lifted search tree in the reference's draw order (k_tree.hip) ------------------
#define TB_STACK 128
#define TB_ERR_NODES 1
#define TB_ERR_STREAM 2
#define TB_ERR_STACK 3
struct TreeBuildParams {
    const float* x;
    uint32_t n, d, leaf, max_depth;
    const uint32_t* stream;   // MT19937 outputs in order
    uint64_t stream_len;
    uint32_t* perm;           // n: point order, leaves contiguous
    uint32_t* tmp;            // n
    uint8_t* side;            // n
    uint32_t* left;           // node_cap each (pre-order ids, root 0)
    uint32_t* right;
    float* offset;
    float* hyper;             // node_cap * d
    uint32_t* leaf_start;
    uint32_t* leaf_size;
    uint32_t node_cap;
    uint32_t* status;         // nodes, error, draws lo, draws hi
};
hipError_t launch_tree_build(const TreeBuildParams& P, hipStream_t s);
hipError_t launch_tree_entries(const uint32_t* left, const uint32_t* leaf_start, const uint32_t* leaf_size,
                               const uint32_t* perm, uint32_t nodes, uint32_t* entry, hipStream_t s);

// ---- relaxed LRQ objective, forward + backward (k_lrq_relax.hip) ----------------------
struct RelaxParams {
    const float* cut;      // n x nk
    uint32_t n, nk, rankm, ncode;
    int mode;              // 0 softmax (nearest), 1 sigmoid (floor)
    const float* mu;       // nk
    const float* G;        // rankm x nk
    const float* cb;       // nk x ncode
    float inv_tau, gscale;
    float *c, *gc;         // n x rankm (work)
    float *e, *q, *mx, *sm, *ge;   // n x nk (work; q holds z = e - q)
    float* loss_rows;      // n
    float* gG;             // rankm x nk (nullable, accumulated)
    float* gcb;            // nk x ncode (nullable, accumulated)
};
hipError_t launch_relax(const RelaxParams& P, hipStream_t s);

// ---- device-wide primitives (k_prims.hip, rocPRIM) ---------------------------------
hipError_t prim_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint64_t m, void* tmp,
                                   size_t& tmp_bytes, hipStream_t s);
hipError_t prim_sort_pairs_u64_u32(const uint64_t* kin, uint64_t* kout, const uint32_t* vin,
                                   uint32_t* vout, uint64_t m, int end_bit, void* tmp,
                                   size_t& tmp_bytes, hipStream_t s);
