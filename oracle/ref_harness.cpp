// ref_harness.cpp -- exposes the reference's own compiled functions through a
// C interface so the tests can pin the oracle (oracle/anqi_oracle.c) against
// them.  TEST INFRASTRUCTURE ONLY; built by oracle/Makefile (target "ref")
// from the reference sources where they lie, into oracle/_ref/.  It includes
// the reference's src/rknn_query.cpp unchanged (its main() renamed) so that
// CompactLrqVerifier and the helpers in that file are the reference's code.
#define main anqi_reference_rknn_query_main
#include "src/rknn_query.cpp"
#undef main

#include <omp.h>
#include <thread>

using nndgraph::HeapList;
using nndgraph::Neighbor;
using nndgraph::NeighborPriorityQueue;
using nndgraph::Parms;
using nndgraph::nndindex;

extern "C" {

float ref_ip(const float* x, const float* y, uint32_t d) {
    return DistCal::InnerProductSIMD16ExtAVX512_(x, y, d);
}
float ref_ip_free(const float* x, const float* y, uint32_t d) {
    return InnerProductSIMD16ExtAVX512_(x, y, d);
}
float ref_l2(const float* x, const float* y, uint32_t d) {
    return DistCal::L2SqrSIMD16ExtAVX512__(x, y, d);
}
float ref_norm(const float* x, uint32_t d) {
    std::vector<float> tmp;
    float* a = static_cast<float*>(_mm_malloc(((d + 15) / 16) * 16 * sizeof(float) + 64, 64));
    std::memcpy(a, x, d * sizeof(float));
    float r = DistCal::NormSIMD16ExtAVX512__(a, d);
    _mm_free(a);
    return r;
}
// The postfilter key lambda of rknn_query.cpp:2555-2561 (same expression,
// compiled with the reference's flags).
float ref_recheck_key(const float* o, const float* q, uint32_t d) {
    float ip = DistCal::InnerProductSIMD16ExtAVX512_(o, q, d);
    float on = DistCal::InnerProductSIMD16ExtAVX512_(o, o, d);
    float qn = DistCal::InnerProductSIMD16ExtAVX512_(q, q, d);
    return on + qn - 2.0f * ip;
}

// The "ip" metric: nndindex's distance (DistCal bound to "ip",
// include/distance.h:170-186), the postfilter key -<o, q> of
// rknn_query.cpp:2555-2561, and the query transform with its R2 bound of
// rknn_query.cpp:2719-2730 (same expressions, compiled with the reference's
// flags; the reference computes them inline in main).
float ref_ipdist(const float* x, const float* y, uint32_t d) {
    DistCal dc(d, "ip");
    return dc.calculate(x, y);
}
float ref_recheck_key_ip(const float* o, const float* q, uint32_t d) {
    float ip = DistCal::InnerProductSIMD16ExtAVX512_(o, q, d);
    return -ip;
}
double ref_query_lift(const float* q, uint32_t d, double scale, double Mstar2, int is_ip, float* qbar_t) {
    double qn2s = 0;
    for (size_t j = 0; j < d; j++) { double v = (double)q[j] * scale; qbar_t[j] = (float)v; qn2s += v * v; }
    qbar_t[d] = 1.0f;
    qbar_t[d + 1] = 0.0f;
    return is_ip ? (qn2s + 1.0 + Mstar2) : (Mstar2 + 1.0);
}

// HeapList::push_nolock over a stream; returns accepted count.
int ref_heap_stream(uint32_t cap, const uint32_t* nodes, const float* vals, uint32_t m,
                    uint32_t* out_idx, float* out_val) {
    HeapList h(1, cap, false);
    int acc = 0;
    for (uint32_t t = 0; t < m; t++) acc += h.push_nolock(0, nodes[t], vals[t]) ? 1 : 0;
    for (uint32_t s = 0; s < cap; s++) { out_idx[s] = h.indices[0][s]; out_val[s] = h.values[0][s]; }
    return acc;
}

// NeighborPriorityQueue: ops[t] == 0 -> insert (ids[t], d[t]); 1 -> pop.
// popped ids written to popped; final contents to out_id/out_d; returns size.
uint32_t ref_queue_stream(uint32_t cap, const uint8_t* ops, const uint32_t* ids,
                          const float* d, uint32_t m, uint32_t* popped,
                          uint32_t* out_id, float* out_d, uint32_t* out_cur) {
    NeighborPriorityQueue q(cap);
    uint32_t np = 0;
    for (uint32_t t = 0; t < m; t++) {
        if (ops[t] == 0) q.insert(Neighbor(ids[t], d[t]));
        else if (q.has_unexpanded_node()) popped[np++] = q.closest_unexpanded().id();
    }
    for (size_t s = 0; s < q.size(); s++) { out_id[s] = q[s].pid; out_d[s] = q[s].distance; }
    size_t cur = 0;
    while (cur < q.size() && q[cur].expanded()) cur++;
    *out_cur = (uint32_t)cur;
    return (uint32_t)q.size();
}

// ---- nndindex over an fbin file ------------------------------------------
struct RefIndex {
    nndindex* idx;
};

void* ref_index_open(const char* fbin, uint64_t n, uint64_t dim, uint64_t M,
                     uint64_t nn_k, uint64_t explore, float ef_alpha, float alpha,
                     const char* folder) {
    Parms p;
    p.n = n; p.dim = dim; p.meric = "l2"; p.random_seed = 10;
    p.nn_k = nn_k; p.M = M; p.leaf_size = 100; p.max_depth = 100; p.n_trees = 1;
    p.ef_alpha = ef_alpha; p.prune_alpha = alpha; p.explore_range = explore;
    p.n_threads = 1; p.folder_path = folder;
    RefIndex* r = new RefIndex;
    r->idx = new nndindex(fbin, p);
    return r;
}
void ref_index_close(void* h) {
    RefIndex* r = static_cast<RefIndex*>(h);
    delete r->idx;
    delete r;
}
float ref_index_d2(void* h, uint64_t a, uint64_t b) {
    return static_cast<RefIndex*>(h)->idx->d2(a, b);
}
float ref_index_norm(void* h, uint64_t a) {
    return static_cast<RefIndex*>(h)->idx->vnorm(a);
}

// Reads a graph cache file in the reference's save_index format
// (include/index.h:1810-1915) into the index.  Returns 0 on success.
int ref_index_load_graph(void* h, const char* path) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    std::ifstream f(path, std::ios::binary);
    if (!f) return 1;
    size_t fn = 0, fdim = 0, fM = 0;
    float ct = 0;
    f.read((char*)&fn, sizeof(fn)); f.read((char*)&fdim, sizeof(fdim));
    f.read((char*)&fM, sizeof(fM)); f.read((char*)&ct, sizeof(ct));
    if (!f || fn != ix->n || fdim != ix->dim || fM != ix->M) return 2;
    for (size_t i = 0; i < ix->n; i++) {
        uint32_t d = 0;
        if (!f.read((char*)&d, 4) || d > ix->M || !f.read((char*)ix->lnk(i), d * 4)) return 3;
        ix->deg(i) = d;
    }
    return ix->load_tree(f) ? 0 : 4;
}

uint32_t ref_tree_entry(void* h, const float* q) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    std::vector<float> qq(q, q + ix->dim);
    return ix->search_tree.get_leave_mediod(qq.data(), ix->dist, ix->random_seed);
}

uint32_t ref_range_search(void* h, const float* q, double R2, uint64_t L,
                          uint32_t* out, uint32_t cap, uint64_t* ndist) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    std::vector<float> qq(q, q + ix->dim);
    std::vector<uint16_t> mark(ix->n, 0);
    std::vector<uint32_t> res;
    size_t nd = 0;
    ix->range_search(qq.data(), R2, 0.0f, L, res, nd, mark.data(), 1, nullptr, 0.5f, false);
    for (size_t t = 0; t < res.size() && t < cap; t++) out[t] = res[t];
    *ndist = nd;
    return (uint32_t)res.size();
}

void ref_graph_arrays(void* h, uint32_t* deg, uint32_t* links) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    for (size_t i = 0; i < ix->n; i++) {
        deg[i] = ix->deg(i);
        for (uint32_t j = 0; j < deg[i]; j++) links[i * ix->M + j] = ix->lnk(i)[j];
    }
}

static HeapList* heap_from(nndindex* ix, const uint32_t* hidx, const float* hval, uint32_t cap) {
    HeapList* g = new HeapList(ix->n, cap, false);
    for (size_t i = 0; i < ix->n; i++)
        for (uint32_t s = 0; s < cap; s++) {
            g->indices[i][s] = hidx[i * cap + s];
            g->values[i][s] = hval[i * cap + s];
        }
    return g;
}
static void heap_to(const HeapList* g, uint32_t* hidx, float* hval, uint32_t cap) {
    for (size_t i = 0; i < g->n_heaps; i++)
        for (uint32_t s = 0; s < cap; s++) {
            hidx[i * cap + s] = g->indices[i][s];
            hval[i * cap + s] = g->values[i][s];
        }
}

// build_index warm start (index.h:1204-1218) with the given flat ids.
void ref_warm_start(void* h, const uint32_t* ids, uint32_t width, uint32_t cap,
                    uint32_t* hidx, float* hval) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    HeapList g(ix->n, cap, false);
    for (size_t i = 0; i < ix->n; i++)
        for (size_t k = 0; k < width; k++) {
            uint32_t j = ids[i * width + k];
            if (j != (uint32_t)i && j < ix->n) g.push_nolock(i, j, ix->d2(i, j));
        }
    heap_to(&g, hidx, hval, cap);
}

// nndescent_refine with one thread (index.h:1468-1487).
void ref_nndescent(void* h, uint32_t* hidx, float* hval, uint32_t cap, int iters) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    HeapList* g = heap_from(ix, hidx, hval, cap);
    ix->n_threads = 1;
    {
        nndgraph::performance_recorder rec("harness nndescent");
        setenv("ANQI_GRAPHNND", std::to_string(iters).c_str(), 1);
        ix->nndescent_refine(g, rec, iters);
    }
    heap_to(g, hidx, hval, cap);
    delete g;
}

// push_reverse (1 thread) + gather_candidates for row i.  Returns pool size.
uint32_t ref_gather(void* h, const uint32_t* hidx, const float* hval, uint32_t cap,
                    uint32_t i, uint64_t pool_cap, float* pd, uint32_t* pi, uint32_t pmax) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    HeapList* g = heap_from(ix, hidx, hval, cap);
    g->push_reverse(1);
    std::vector<std::pair<float, uint32_t>> pool;
    ix->gather_candidates(i, g, pool, pool_cap);
    for (size_t t = 0; t < pool.size() && t < pmax; t++) { pd[t] = pool[t].first; pi[t] = pool[t].second; }
    delete g;
    return (uint32_t)pool.size();
}

uint32_t ref_prune(void* h, uint32_t i, const float* pd, const uint32_t* pi, uint32_t m,
                   uint32_t* out) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    ix->configure_edge_residual_cover();
    std::vector<std::pair<float, uint32_t>> pool;
    for (uint32_t t = 0; t < m; t++) pool.emplace_back(pd[t], pi[t]);
    std::vector<uint32_t> ids;
    ix->prune_pool_to_ids(i, pool, ids);
    for (size_t t = 0; t < ids.size(); t++) out[t] = ids[t];
    return (uint32_t)ids.size();
}

void ref_commit(void* h, const uint32_t* O, const uint32_t* ocnt, uint32_t M,
                uint32_t* links, uint32_t* deg) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    ix->configure_edge_residual_cover();
    ix->n_threads = 1;
    std::vector<std::vector<uint32_t>> OO(ix->n);
    for (size_t i = 0; i < ix->n; i++) OO[i].assign(O + i * M, O + i * M + ocnt[i]);
    ix->commit_edges(OO);
    ref_graph_arrays(h, deg, links);
}

// build_index (index.h:1071-1295) on one thread under the current environment
// (ANQI_REUSE_ORIG_KNN, ANQI_VAMANA*, ...), warm-started from the given ids:
// the search tree, graph and Vamana passes as the reference builds them.
// Exports the graph and graph_entry_.  The build runs on a fresh thread: the
// tree's generator is a thread_local engine seeded by its first call
// (include/utils.h:169-173), so only a new thread draws the tree of the
// index's own seed.
void ref_build_graph(void* h, const uint32_t* ids, uint32_t width, uint32_t* deg, uint32_t* links,
                     uint32_t* graph_entry) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    ix->n_threads = 1;
    std::vector<std::vector<uint32_t>> nbr(ix->n);
    for (size_t i = 0; i < ix->n; i++) nbr[i].assign(ids + i * width, ids + (i + 1) * width);
    std::thread t([&]() { ix->build_index(true, false, &nbr); });
    t.join();
    ref_graph_arrays(h, deg, links);
    *graph_entry = ix->graph_entry_;
}

// greedy_collect (index.h:831-855; entry = UINT32_MAX) or
// greedy_collect_from_entry (:857-880) for row i over the index's current
// graph and tree.  Returns the pool size (pool truncated at pmax).
uint32_t ref_vamana_collect(void* h, uint32_t i, uint32_t entry, uint64_t Lb, float* pd, uint32_t* pi,
                            uint32_t pmax) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    std::vector<std::pair<float, uint32_t>> pool;
    if (entry == UINT32_MAX) ix->greedy_collect(i, pool, Lb);
    else ix->greedy_collect_from_entry(i, pool, Lb, entry);
    for (size_t t = 0; t < pool.size() && t < pmax; t++) { pd[t] = pool[t].first; pi[t] = pool[t].second; }
    return (uint32_t)pool.size();
}

// std::shuffle of 0..n-1 with std::mt19937_64(seed), as vamana_batch_refine
// draws its insertion order (index.h:956-958).
void ref_shuffle(uint32_t n, uint64_t seed, uint32_t* out) {
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::mt19937_64 rng(seed);
    std::shuffle(order.begin(), order.end(), rng);
    std::copy(order.begin(), order.end(), out);
}

// choose_vamana_entry (index.h:190-222) under the current environment.
uint32_t ref_vamana_entry(void* h) {
    return static_cast<RefIndex*>(h)->idx->choose_vamana_entry();
}

// ---- CompactLrqVerifier (rknn_query.cpp:603-839), policy "floor" ---------
void ref_lrq_build(uint64_t n, uint32_t nk, uint32_t rankm, const float* mu,
                   const float* basis, const float* codebook, uint32_t code_size,
                   const float* cutoffs, float* coeff, uint8_t* codes) {
    omp_set_num_threads(1);
    CompactLrqVerifier v;
    std::vector<float> m(mu, mu + nk), b(basis, basis + (size_t)rankm * nk),
        cb(codebook, codebook + (size_t)nk * code_size);
    v.configure(n, nk, (int)rankm, 8, "floor", m, b, cb, code_size);
    std::vector<float> src(cutoffs, cutoffs + n * nk);
    v.build(src);
    std::memcpy(coeff, v.coeff.data(), v.coeff.size() * sizeof(float));
    std::memcpy(codes, v.packed_codes.data(), v.packed_codes.size());
}

void ref_lrq_materialize_metric(uint64_t n, uint32_t nk, uint32_t rankm, const float* mu,
                                const float* basis, const float* codebook, uint32_t code_size,
                                const float* coeff, const uint8_t* codes, uint32_t k,
                                double slack, int is_ip, float* thr) {
    CompactLrqVerifier v;
    std::vector<float> m(mu, mu + nk), b(basis, basis + (size_t)rankm * nk),
        cb(codebook, codebook + (size_t)nk * code_size);
    v.configure(n, nk, (int)rankm, 8, "floor", m, b, cb, code_size);
    v.coeff.assign(coeff, coeff + n * rankm);
    v.packed_codes.assign(codes, codes + n * nk);
    std::vector<float> t;
    v.materialize_k(k, slack, is_ip != 0, t);
    std::memcpy(thr, t.data(), n * sizeof(float));
}

void ref_lrq_materialize(uint64_t n, uint32_t nk, uint32_t rankm, const float* mu,
                         const float* basis, const float* codebook, uint32_t code_size,
                         const float* coeff, const uint8_t* codes, uint32_t k,
                         double slack, float* thr) {
    CompactLrqVerifier v;
    std::vector<float> m(mu, mu + nk), b(basis, basis + (size_t)rankm * nk),
        cb(codebook, codebook + (size_t)nk * code_size);
    v.configure(n, nk, (int)rankm, 8, "floor", m, b, cb, code_size);
    v.coeff.assign(coeff, coeff + n * rankm);
    v.packed_codes.assign(codes, codes + n * nk);
    std::vector<float> t;
    v.materialize_k(k, slack, false, t);
    std::memcpy(thr, t.data(), n * sizeof(float));
}

uint32_t ref_lrq_encode_value(double v, const float* centers, uint32_t code_size) {
    return encode_learned_residual(v, centers, code_size, "floor");
}

// The reference's build_rptree (rptrees.h:155-163), run on a thread of its own
// so that generate_random_integer's thread_local engine starts from `seed`,
// as it does for the lifted search tree (index.h:1116 builds one tree inside a
// parallel region on a worker whose engine is still unseeded).
void* ref_tree_build(const float* data, uint64_t n, uint64_t dim, uint64_t leaf, uint64_t max_depth,
                     uint64_t seed) {
    auto* t = new nndgraph::rptree();
    std::thread th([&] {
        *t = nndgraph::build_rptree(const_cast<float*>(data), n, dim, leaf, false, max_depth, seed);
    });
    th.join();
    return t;
}

uint64_t ref_tree_nodes(void* h, uint64_t* total_idx) {
    auto* t = static_cast<nndgraph::rptree*>(h);
    uint64_t s = 0;
    for (const auto& nd : t->nodes) s += nd.indices.size();
    *total_idx = s;
    return t->nodes.size();
}

void ref_tree_export(void* h, uint64_t* left, uint64_t* right, float* offset, float* hyper, uint64_t dim,
                     uint64_t* idx_off, uint32_t* idx) {
    auto* t = static_cast<nndgraph::rptree*>(h);
    idx_off[0] = 0;
    for (size_t k = 0; k < t->nodes.size(); k++) {
        const auto& nd = t->nodes[k];
        left[k] = nd.left;
        right[k] = nd.right;
        offset[k] = nd.offset;
        for (uint64_t j = 0; j < dim; j++) hyper[k * dim + j] = nd.hyperplane.empty() ? 0.f : nd.hyperplane[j];
        std::memcpy(idx + idx_off[k], nd.indices.data(), 4 * nd.indices.size());
        idx_off[k + 1] = idx_off[k] + nd.indices.size();
    }
}

void ref_tree_free(void* h) { delete static_cast<nndgraph::rptree*>(h); }

}  // extern "C"

// ---- CPU baseline: the reference's query loop body (rknn_query.cpp:2705-2776)
// over an index loaded into the harness: lift q, range_search, postfilter
// into an unordered_set.  Ground-truth evaluation (also inside the
// reference's timer) is not included.  Returns elapsed seconds.
extern "C" double ref_rknn_batch(void* h, const float* base, uint32_t d, const float* q, uint32_t nq,
                                 double scale, double R2, uint64_t L, const float* thr, int threads,
                                 uint32_t repeats, uint64_t* tot_cand, uint64_t* tot_pred,
                                 uint64_t* tot_dist) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    const size_t n = ix->n, dl = ix->dim;
    omp_set_num_threads(threads);
    uint64_t c = 0, p = 0, dd = 0;
    auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel reduction(+ : c, p, dd)
    {
        std::vector<uint16_t> mark(n, 0);
        uint16_t ver = 0;
        std::vector<uint32_t> cand;
        std::vector<float> qb(dl);
#pragma omp for schedule(dynamic, 16)
        for (size_t task = 0; task < (size_t)nq * repeats; task++) {
            const size_t i = task % nq;
            const float* qq = q + i * d;
            for (size_t j = 0; j < d; j++) qb[j] = (float)((double)qq[j] * scale);
            qb[d] = 1.0f;
            qb[d + 1] = 0.0f;
            if (++ver == 0) { std::fill(mark.begin(), mark.end(), 0); ver = 1; }
            size_t nd = 0;
            ix->range_search(qb.data(), R2, 0.0f, L, cand, nd, mark.data(), ver, nullptr, 0.5f, false);
            std::unordered_set<uint32_t> pred;
            for (uint32_t id : cand) {
                const float* o = base + (size_t)id * d;
                float ip = DistCal::InnerProductSIMD16ExtAVX512_(o, qq, d);
                float on = DistCal::InnerProductSIMD16ExtAVX512_(o, o, d);
                float qn = DistCal::InnerProductSIMD16ExtAVX512_(qq, qq, d);
                if (on + qn - 2.0f * ip <= (double)thr[id]) pred.insert(id);
            }
            c += cand.size();
            p += pred.size();
            dd += nd + cand.size();
        }
    }
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    *tot_cand = c;
    *tot_pred = p;
    *tot_dist = dd;
    return secs;
}

// Per-query form of ref_rknn_batch (one pass): candidate, accepted and
// distance counts, the entry point, and the number of |margin| < EPS ties met
// on the search-tree descent (where the reference draws a random side from
// the calling thread's engine, rptrees.h:116-118).
extern "C" void ref_rknn_detail(void* h, const float* base, uint32_t d, const float* q, uint32_t nq,
                                double scale, double R2, uint64_t L, const float* thr, int threads,
                                uint32_t* cnt, uint32_t* pred, uint32_t* ndist, uint32_t* entry,
                                uint32_t* ties) {
    nndindex* ix = static_cast<RefIndex*>(h)->idx;
    const size_t n = ix->n, dl = ix->dim;
    omp_set_num_threads(threads);
#pragma omp parallel
    {
        std::vector<uint16_t> mark(n, 0);
        uint16_t ver = 0;
        std::vector<uint32_t> cand;
        std::vector<float> qb(dl);
#pragma omp for schedule(dynamic, 16)
        for (size_t i = 0; i < nq; i++) {
            const float* qq = q + i * d;
            for (size_t j = 0; j < d; j++) qb[j] = (float)((double)qq[j] * scale);
            qb[d] = 1.0f;
            qb[d + 1] = 0.0f;
            uint32_t nt = 0;
            {
                const auto& T = ix->search_tree;
                size_t k = T.nodes.size() - 1;
                while (T.nodes[k].left != UINT32_MAX) {
                    float m = DistCal::InnerProductSIMD16ExtAVX512_(T.nodes[k].hyperplane.data(), qb.data(),
                                                                    (uint32_t)dl) - T.nodes[k].offset;
                    if (m < -EPS) k = T.nodes[k].left;
                    else if (m > EPS) k = T.nodes[k].right;
                    else { nt++; k = T.nodes[k].left; }
                }
                entry[i] = T.nodes[k].indices.empty() ? 0 : T.nodes[k].indices[0];
            }
            ties[i] = nt;
            if (++ver == 0) { std::fill(mark.begin(), mark.end(), 0); ver = 1; }
            size_t nd = 0;
            ix->range_search(qb.data(), R2, 0.0f, L, cand, nd, mark.data(), ver, nullptr, 0.5f, false);
            std::unordered_set<uint32_t> pr;
            for (uint32_t id : cand) {
                const float* o = base + (size_t)id * d;
                float ip = DistCal::InnerProductSIMD16ExtAVX512_(o, qq, d);
                float on = DistCal::InnerProductSIMD16ExtAVX512_(o, o, d);
                float qn = DistCal::InnerProductSIMD16ExtAVX512_(qq, qq, d);
                if (on + qn - 2.0f * ip <= (double)thr[id]) pr.insert(id);
            }
            cnt[i] = (uint32_t)cand.size();
            pred[i] = (uint32_t)pr.size();
            ndist[i] = (uint32_t)nd;
        }
    }
}
