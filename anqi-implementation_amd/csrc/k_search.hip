// k_search.hip -- RkNN range search over the lifted graph plus the fused
// verifier postfilter.  One wavefront per query.
//
// Semantics follow nndindex::range_search (include/index.h:1535-1652) in the
// pinned profile (CENTER score, radj == nullptr, eps = 0): a beam of width L
// keyed on lifted L2^2 from the search-tree entry, then an unbounded flood over
// every node with D <= R2; collected candidates are the nodes with D <= R2 in
// discovery order.  The comparison double(float D) <= R2 is done as
// D <= R2f where R2f is the largest float <= R2: R2 = M*^2 + 1 for L2
// (computed on the host), and per query R2 = |s q|^2 + 1 + M*^2 for the
// inner product (rknn_query.cpp:2728-2729, k_query_r2 below).
// The postfilter is src/rknn_query.cpp:2762-2766: key = on + qn - 2 ip in base
// space (each term by the FP32 contract), or -ip for the inner product
// (:2555-2561), accept iff key <= t_k[o].
//
// Batch semantics: merging one expansion's (up to 64) neighbours into the
// sorted beam at once equals inserting them one by one (total (dist, id)
// order, ids unique thanks to the visited set).  The flood is a closure, so
// its visiting order does not change the collected set; it is processed in
// FIFO order anyway, so candidate lists match the reference's order.
#include <cstring>

#include "common.hpp"
#include "kernels.hpp"

// Flood: up to FL FIFO nodes are expanded together (their adjacency rows,
// then all their new neighbours' rows, are fetched in one round each), as
// long as their neighbours fit the SR_TID-entry id buffer.  FL is
// SR_FLOOD_PLAIN in the plain form and SR_FLOOD_MAX in the persistent form.
#define SR_FLOOD_PLAIN 4
#define SR_FLOOD_MAX 16
#define SR_TID 128

namespace {

struct SearchLds {
    float* bd0;        // beam distances (double buffer 0 / 1)
    float* bd1;
    uint32_t* bp0;     // beam ids, bit31 = expanded
    uint32_t* bp1;
    // buffer selection by value, so the struct stays in registers
    __device__ float* bd(int b) const { return b ? bd1 : bd0; }
    __device__ uint32_t* bp(int b) const { return b ? bp1 : bp0; }
    uint32_t* hash;    // visited set (LDS tiers)
    uint32_t* gvis;    // visited bitmap of n bits (global tier)
    uint32_t* res;     // collected ids (FIFO), first rcap entries
    uint32_t* spill;   // the query's output row: FIFO entries from rcap on
    uint32_t* tid;     // SR_TID new ids
    float* tdv;        // SR_TID distances
    uint64_t* cs;      // 64 sorted candidate keys
    float r2;          // the query's R2f (collect and flood iff D <= r2)
};

// Collected-candidate FIFO: entries below rcap live in LDS, later ones in
// the query's output row, at the position the final list gives them anyway
// (the post-filter rewrites out[j] = fifo[j]).  Only more than
// max(rcap, out_cap) candidates overflow.
__device__ __forceinline__ uint32_t fifo_cap(const SearchParams& P) { return max(P.rcap, P.out_cap); }
__device__ __forceinline__ void fifo_put(const SearchParams& P, const SearchLds& S, uint32_t j, uint32_t v) {
    if (j < P.rcap) S.res[j] = v; else S.spill[j] = v;
}
__device__ __forceinline__ uint32_t fifo_get(const SearchParams& P, const SearchLds& S, uint32_t j) {
    return j < P.rcap ? S.res[j] : S.spill[j];
}

// GV: global tier -- the visited set and the FIFO live in global memory
// (set by the kernel), LDS holds the beam and the id/distance buffers.
template <bool GV>
__device__ __forceinline__ SearchLds carve(char* base, const SearchParams& P) {
    SearchLds s;
    size_t o = 0;
    auto take = [&](size_t bytes) { char* p = base + o; o += (bytes + 15) & ~size_t(15); return p; };
    if (P.rbeam) {
        // register beam: one (dist, id) key exchange area of Lmax entries at bd0
        s.bd0 = (float*)take(sizeof(uint64_t) * P.Lmax);
        s.bd1 = s.bd0; s.bp0 = nullptr; s.bp1 = nullptr; s.cs = nullptr;
    } else {
        s.bd0 = (float*)take(sizeof(float) * P.Lmax);
        s.bd1 = (float*)take(sizeof(float) * P.Lmax);
        s.bp0 = (uint32_t*)take(sizeof(uint32_t) * P.Lmax);
        s.bp1 = (uint32_t*)take(sizeof(uint32_t) * P.Lmax);
        s.cs = (uint64_t*)take(sizeof(uint64_t) * 64);
    }
    s.gvis = nullptr;
    if constexpr (GV) {
        s.hash = nullptr;
        s.res = nullptr;
    } else {
        s.hash = (uint32_t*)take(sizeof(uint32_t) * (P.hmask + 1));
        s.res = (uint32_t*)take(sizeof(uint32_t) * P.rcap);
    }
    s.tid = (uint32_t*)take(sizeof(uint32_t) * SR_TID);
    s.tdv = (float*)take(sizeof(float) * SR_TID);
    return s;
}

// Visited-set insert: 1 if id was not yet visited, 0 if it was, -1 if the
// LDS set has no room for it (LDS tiers only; the bitmap never fills).
template <bool GV>
__device__ __forceinline__ int vis_insert(const SearchParams& P, const SearchLds& S, uint32_t id) {
    if constexpr (GV) {
        const uint32_t bit = 1u << (id & 31u);
        return (atomicOr(S.gvis + (id >> 5), bit) & bit) ? 0 : 1;
    } else {
        return hset_insert(S.hash, P.hmask, id);
    }
}

// Node c's adjacency row, one neighbour per lane (U32_NONE on lanes >= M).
// Lanes below dg may hold neighbours: dg is the degree, or M with the
// sentinel rows (P.plinks: U32_NONE past the degree, so no degree read).  A
// lane holds a neighbour iff lane < dg and its id is not U32_NONE.  PK: the
// sentinel rows (P.plinks, set exactly when the kernel is instantiated with PK).
template <bool PK>
__device__ __forceinline__ void adj_load(const SearchParams& P, uint32_t c, uint32_t& dg, uint32_t& nbr) {
    const uint32_t lane = (uint32_t)__lane_id();
    const uint32_t* row = (PK ? P.plinks : P.links) + (size_t)c * P.M;
    nbr = lane < P.M ? row[lane] : U32_NONE;
    dg = PK ? P.M : P.deg[c];
}

// Expands node c: visits its unvisited neighbours, computes D, appends
// collected ones to the FIFO.  In beam mode also merges them into the beam.
// Returns false on a capacity overflow.
// ---- byte rows -----------------------------------------------------------
// When every base value is one of the floats 0..255, a row is kept as d bytes
// (P.brows, d = 16 MC), ordered for the 16-lane groups: lane l's MC bytes
// l MC .. l MC + MC - 1 are elements l, 16 + l, ..., 16 (MC - 1) + l.  The
// base value is float(b).  The lifted main coordinate is the lift's
// float(double(b) * scale) (rknn_query.cpp:2203-2236); it is rebuilt in FP32
// as fma(b, s_hi, b * s_lo) with s_hi + s_lo = scale to about 2^-48, which
// is within 2^-47 of b * scale and so rounds the same way unless b * scale
// lies that close to a rounding boundary.  anqi_byte_rows checks the rebuilt
// value against the lifted rows on every element before an index uses the
// bytes, so a base with such a value keeps its float rows.
template <int MC>
__device__ __forceinline__ void brow_fetch(const uint8_t* __restrict__ B, uint32_t id, int l, uint32_t (&w)[2]) {
    const uint8_t* p = B + (size_t)id * (16 * MC) + (size_t)l * MC;
    if constexpr (MC == 8) {
        const uint2 v = *reinterpret_cast<const uint2*>(p);
        w[0] = v.x;
        w[1] = v.y;
    } else if constexpr (MC % 2 == 0) {
        const uint16_t* h = reinterpret_cast<const uint16_t*>(p);
        w[0] = 0u;
        w[1] = 0u;
#pragma unroll
        for (int i = 0; i < MC / 2; i++) w[i >> 1] |= (uint32_t)h[i] << (16 * (i & 1));
    } else {
        w[0] = 0u;
        w[1] = 0u;
#pragma unroll
        for (int i = 0; i < MC; i++) w[i >> 2] |= (uint32_t)p[i] << (8 * (i & 3));
    }
}
__device__ __forceinline__ uint32_t brow_byte(const uint32_t (&w)[2], int t) {
    return (w[t >> 2] >> (8 * (t & 3))) & 255u;
}
// scale split into s_hi = float(scale) and s_lo = float(scale - s_hi)
struct BScale {
    float hi, lo;
};
__device__ __forceinline__ BScale bscale(double scale) {
    const float hi = (float)scale;
    return BScale{hi, (float)(scale - (double)hi)};
}
// the lifted coordinate of byte value b
__device__ __forceinline__ float brow_lifted(uint32_t b, BScale s) {
    const float bf = (float)b;
    return fmaf(bf, s.hi, bf * s.lo);
}

// rows_batch MODE 2 / 3 (split lifted rows) over byte rows: the same FMA
// order and reduction tree, the coordinates rebuilt from the bytes.
template <int MC, int RB>
__device__ __forceinline__ void bytes_pass(const QRow& q, const uint8_t* __restrict__ B,
                                           const float* __restrict__ ltail, BScale scale, const uint32_t* ids, int k,
                                           int l, int g, float* out, int s0) {
    uint32_t w[RB][2];
    float nb[RB];
#pragma unroll
    for (int j = 0; j < RB; j++) {
        const int s = s0 + g + 4 * j;
        w[j][0] = 0u;
        w[j][1] = 0u;
        nb[j] = 0.0f;
        if (s < k) {
            const uint32_t id = ids[s];
            brow_fetch<MC>(B, id, l, w[j]);
            nb[j] = ltail[id];
        }
    }
#pragma unroll
    for (int j = 0; j < RB; j++) {
        float acc = 0.0f;
#pragma unroll
        for (int t = 0; t < MC; t++) {
            const float a = q.v[t] - brow_lifted(brow_byte(w[j], t), scale);
            acc = fmaf(a, a, acc);
        }
        const float res = tree16(acc) + nb[j];
        const int s = s0 + g + 4 * j;
        if (l == 0 && s < k) out[s] = res;
    }
}

template <int MC, int RB>
__device__ __forceinline__ void bytes_tail(const QRow& q, const uint8_t* __restrict__ B,
                                           const float* __restrict__ ltail, BScale scale, const uint32_t* ids, int k,
                                           int l, int g, float* out, int s0) {
    if constexpr (RB > 1) {
        if (k - s0 <= 2 * RB) {
            bytes_tail<MC, RB / 2>(q, B, ltail, scale, ids, k, l, g, out, s0);
            return;
        }
    }
    bytes_pass<MC, RB>(q, B, ltail, scale, ids, k, l, g, out, s0);
}

template <int MC, int RB = 4>
__device__ __forceinline__ void rows_batch_bytes(const QRow& q, const uint8_t* __restrict__ B,
                                                 const float* __restrict__ ltail, BScale scale,
                                                 const uint32_t* ids, int k, int l, int g, float* out) {
    int s0 = 0;
    for (; k - s0 > 4 * RB; s0 += 4 * RB) bytes_pass<MC, RB>(q, B, ltail, scale, ids, k, l, g, out, s0);
    if (s0 < k) bytes_tail<MC, RB>(q, B, ltail, scale, ids, k, l, g, out, s0);
}

// D for the k new neighbours tid[0..k).  Group g (16 lanes) takes rows g,
// g+4, g+8, g+12 of each 16-row pass and loads all four before reducing any,
// so their fetches overlap.  MC = number of full 16-float chunks of a lifted
// row (d' = 16 MC + r, r < 16); MC == 0 selects the generic loop.
// Split layout (P.lmain, d = 16 MC): the d leading coordinates come from the
// aligned rows of lmain (or the byte rows) and the two lift coordinates'
// reduced term from ltail; D = tree16(main) + tail, the same two terms the
// whole-row form adds.
template <int MC, bool BR>
__device__ __forceinline__ void batch_l2(const SearchParams& P, const QRow& QL, const uint32_t* tid,
                                         float* tdv, int k, int l, int g) {
    if constexpr (MC > 0 && BR) {
        rows_batch_bytes<MC>(QL, P.brows, P.ltail, bscale(__longlong_as_double((long long)P.scale_bits)), tid, k, l,
                             g, tdv);
    } else if constexpr (MC > 0) {
        if (P.lmain) rows_batch<MC, 2>(QL, P.lmain, P.d, P.ltail, 0.0f, tid, k, l, g, tdv);
        else rows_batch<MC, 0>(QL, P.lifted, P.dl, nullptr, 0.0f, tid, k, l, g, tdv);
    } else {
        for (int s = g; s < k; s += 4) {
            float dv = qrow_l2(QL, P.lifted + (size_t)tid[s] * P.dl, P.dl, l);
            if (l == 0) tdv[s] = dv;
        }
    }
}

template <bool BEAM, int MC, bool GV, bool PK, bool BR>
__device__ __forceinline__ bool expand(const SearchParams& P, SearchLds& S, const QRow& QL,
                                       uint32_t c, uint32_t& rcnt, uint32_t& nd,
                                       int& buf, uint32_t& bsize, uint32_t L, int& ovf) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    // degree and neighbour row fetched together
    uint32_t dg, nbr;
    adj_load<PK>(P, c, dg, nbr);
    uint32_t nb = U32_NONE;
    int fresh = 0;
    if ((uint32_t)lane < dg && nbr != U32_NONE) {
        nb = nbr;
        int r = vis_insert<GV>(P, S, nb);
        if (r < 0) ovf = 1;
        fresh = (r == 1);
    }
    if (__any(ovf)) { ovf = 1; return false; }
    uint64_t fm = ballot(fresh);
    int k = popc64(fm);
    if (fresh) S.tid[lanes_below(fm)] = nb;
    __syncthreads();
    batch_l2<MC, BR>(P, QL, S.tid, S.tdv, k, l, g);
    __syncthreads();
    nd += (uint32_t)k;
    float dv = (lane < k) ? S.tdv[lane] : F32_BIG;
    uint32_t myid = (lane < k) ? S.tid[lane] : U32_NONE;
    int hit = (lane < k) && (dv <= S.r2);
    uint64_t hm = ballot(hit);
    int nh = popc64(hm);
    if (rcnt + (uint32_t)nh > fifo_cap(P)) { ovf = 2; return false; }
    if (hit) fifo_put(P, S, rcnt + lanes_below(hm), myid);
    rcnt += (uint32_t)nh;
    if (BEAM) {
        // candidate keys; drop those that cannot enter a full beam
        uint64_t key = (lane < k) ? pkey(dv, myid) : ~0ull;
        float* bd = S.bd(buf);
        uint32_t* bp = S.bp(buf);
        if (bsize == L && lane < k) {
            uint64_t last = pkey(bd[L - 1], bp[L - 1] & 0x7fffffffu);
            if (!(key < last)) key = ~0ull;
        }
        key = wave_sort64(key);
        uint64_t km = ballot(key != ~0ull);
        int kc = popc64(km);
        if (kc == 0) return true;
        S.cs[lane] = key;
        __syncthreads();
        float* od = S.bd(buf ^ 1);
        uint32_t* op = S.bp(buf ^ 1);
        // beam elements move right by the number of smaller candidates
        for (uint32_t i = lane; i < bsize; i += 64) {
            uint64_t bk = pkey(bd[i], bp[i] & 0x7fffffffu);
            int lo = 0, hi = kc;
            while (lo < hi) { int mid = (lo + hi) >> 1; if (S.cs[mid] < bk) lo = mid + 1; else hi = mid; }
            uint32_t pos = i + (uint32_t)lo;
            if (pos < L) { od[pos] = bd[i]; op[pos] = bp[i]; }
        }
        if (lane < kc) {
            uint32_t lo = 0, hi = bsize;
            while (lo < hi) {
                uint32_t mid = (lo + hi) >> 1;
                if (pkey(bd[mid], bp[mid] & 0x7fffffffu) < key) lo = mid + 1; else hi = mid;
            }
            uint32_t pos = (uint32_t)lane + lo;
            if (pos < L) { od[pos] = pkey_d(key); op[pos] = pkey_id(key); }
        }
        bsize = min(L, bsize + (uint32_t)kc);
        buf ^= 1;
        __syncthreads();
    }
    return true;
}

// Register-resident beam (persistent form): an RKeys list (common.hpp).
// dg, nbr: node c's adjacency row (adj_load), fetched by the caller.
template <int NCH, int MC, bool PK, bool BR>
__device__ __forceinline__ bool expand_rb(const SearchParams& P, SearchLds& S, const QRow& QL, uint32_t dg,
                                          uint32_t nbr, uint32_t& rcnt, uint32_t& nd, RKeys<NCH>& B,
                                          uint32_t& bsize, uint32_t L, int& ovf) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    uint32_t nb = U32_NONE;
    int fresh = 0;
    if ((uint32_t)lane < dg && nbr != U32_NONE) {
        nb = nbr;
        int r = hset_insert(S.hash, P.hmask, nb);
        if (r < 0) ovf = 1;
        fresh = (r == 1);
    }
    if (__any(ovf)) { ovf = 1; return false; }
    const uint64_t fm = ballot(fresh);
    const int k = popc64(fm);
    if (fresh) S.tid[lanes_below(fm)] = nb;
    __syncthreads();
    batch_l2<MC, BR>(P, QL, S.tid, S.tdv, k, l, g);
    __syncthreads();
    nd += (uint32_t)k;
    const float dv = (lane < k) ? S.tdv[lane] : F32_BIG;
    const uint32_t myid = (lane < k) ? S.tid[lane] : U32_NONE;
    const int hit = (lane < k) && (dv <= S.r2);
    const uint64_t hm = ballot(hit);
    const int nh = popc64(hm);
    if (rcnt + (uint32_t)nh > fifo_cap(P)) { ovf = 2; return false; }
    if (hit) fifo_put(P, S, rcnt + lanes_below(hm), myid);
    rcnt += (uint32_t)nh;
    rk_merge<NCH>(B, (lane < k) ? pkey(dv, myid) : ~0ull, bsize, L, (uint64_t*)S.bd0);
    return true;
}

// Flood step over up to nb (<= FL) FIFO entries at head: the
// visited-set inserts run node by node in FIFO order, neighbours in
// adjacency order, so every neighbour is attributed to the same first
// visitor as in a one-node-at-a-time flood; collected nodes are appended in
// that same order.  Their distances are computed in one batch.  The step
// stops before a node whose neighbours might not fit the id buffer; nb is
// set to the number of nodes expanded (at least one: deg <= 64 <= SR_TID).
template <int MC, int FL, bool GV, bool PK, bool BR>
__device__ __forceinline__ bool flood_step(const SearchParams& P, SearchLds& S, const QRow& QL, uint32_t head,
                                           uint32_t& nb, uint32_t& rcnt, uint32_t& nd, int& ovf) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    uint32_t dg[FL], nbr[FL];
#pragma unroll
    for (int t = 0; t < FL; t++) {
        dg[t] = 0;
        nbr[t] = U32_NONE;
        if ((uint32_t)t < nb) adj_load<PK>(P, fifo_get(P, S, head + t), dg[t], nbr[t]);
    }
    int k = 0;
    uint32_t done = nb;
#pragma unroll
    for (int t = 0; t < FL; t++) {
        const int has = (uint32_t)lane < dg[t] && nbr[t] != U32_NONE;
        const int deg = PK ? popc64(ballot(has)) : (int)dg[t];
        if ((uint32_t)t < done && k + deg > SR_TID) done = (uint32_t)t;
        if ((uint32_t)t >= done) continue;
        int fresh = 0;
        if (has) {
            int r = vis_insert<GV>(P, S, nbr[t]);
            if (r < 0) ovf = 1;
            fresh = (r == 1);
        }
        const uint64_t fm = ballot(fresh);
        if (fresh) S.tid[k + lanes_below(fm)] = nbr[t];
        k += popc64(fm);
    }
    nb = done;
    if (__any(ovf)) { ovf = 1; return false; }
    __syncthreads();
    batch_l2<MC, BR>(P, QL, S.tid, S.tdv, k, l, g);
    __syncthreads();
    nd += (uint32_t)k;
    for (int b = 0; b < k; b += 64) {
        const int s = b + lane;
        const int hit = s < k && S.tdv[s] <= S.r2;
        const uint64_t hm = ballot(hit);
        const int nh = popc64(hm);
        if (rcnt + (uint32_t)nh > fifo_cap(P)) { ovf = 2; return false; }
        if (hit) fifo_put(P, S, rcnt + lanes_below(hm), S.tid[s]);
        rcnt += (uint32_t)nh;
    }
    __syncthreads();
    return true;
}

// NCH == 0: beam sorted in LDS (plain form); NCH > 0: register beam of
// 64 NCH positions (persistent form, L <= 64 NCH).  FL: flood step width.
// GV: global tier (NCH == 0 only).
template <int MC, int NCH, int FL, bool GV = false, bool PK = false, bool BR = false>
__device__ void search_one(const SearchParams& P, SearchLds& S, uint32_t qi) {
    const int lane = __lane_id(), l = lane & 15;
    S.spill = P.cand_ids + (size_t)qi * P.out_cap;
    // per-query beam width, clamped to 1..Lmax (the LDS beam areas' size)
    const uint32_t L = max(1u, min(P.Lq ? P.Lq[qi] : P.L, P.Lmax));
    const float* thr = P.thr + (P.kidx ? (size_t)P.kidx[qi] * P.n : 0);
    S.r2 = P.qr2 ? P.qr2[qi] : P.R2f;
    // query transform (rknn_query.cpp:2719-2730), binary64 product rounded to
    // float, then the lift coordinates (1, 0); each lane forms the elements
    // its register row holds (element 16 t + l, and the tail)
    const float* q = P.queries + (size_t)qi * P.d;
    if constexpr (GV) {
        // clear this block's bitmap; the fence orders the stores before the
        // visited-set atomics
        uint4* w = (uint4*)S.gvis;
        for (uint32_t j = lane; j < (P.gstride >> 2); j += 64) w[j] = make_uint4(0u, 0u, 0u, 0u);
        __threadfence();
    } else {
        for (uint32_t j = lane; j <= P.hmask; j += 64) S.hash[j] = 0u;
    }
    QRow QL;
    {
        // lane holds elements lane + 64 k (k < 4, dl <= 256), moved into the
        // register-row layout (element 16 t + l) by shuffles
        float qv0 = 0.0f, qv1 = 0.0f, qv2 = 0.0f, qv3 = 0.0f;
#pragma unroll 1
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t e = (uint32_t)lane + 64u * k;
            float v = 0.0f;
            if (e < P.d) v = sd_to_float(sd_mul(sd_from_float(q[e]), P.scale_bits));
            else if (e == P.d) v = 1.0f;
            if (k == 0) qv0 = v; else if (k == 1) qv1 = v; else if (k == 2) qv2 = v; else qv3 = v;
        }
        auto pick = [&](uint32_t blk) { return blk == 0 ? qv0 : blk == 1 ? qv1 : blk == 2 ? qv2 : qv3; };
        const uint32_t m = P.dl >> 4, r = P.dl & 15;
#pragma unroll
        for (int t = 0; t < QMAX_CHUNKS; t++) {
            const float v = __shfl(pick((uint32_t)t >> 2), (16 * t + l) & 63);
            QL.v[t] = ((uint32_t)t < m) ? v : 0.0f;
        }
        const float tv = __shfl(pick(m >> 2), (int)((16u * m + (uint32_t)l) & 63u));
        QL.tail = ((uint32_t)l < r) ? tv : 0.0f;
    }
    __syncthreads();
    // entry point: search-tree descent (rptrees.h:108-123), ties go left
    uint32_t ep;
    if (P.entries) {
        ep = P.entries[qi];
    } else {
        // a node's children, offset and hyperplane are requested together
        uint32_t node = P.t_root;
        for (;;) {
            const uint32_t lc = P.t_left[node], rc = P.t_right[node];
            const float off = P.t_off[node];
            const float ip = qrow_ip_mc<MC>(QL, P.t_hyper + (size_t)node * P.dl, P.dl, l);
            if (lc == U32_NONE) break;
            const float mg = ip - off;
            node = (mg < -1e-6f) ? lc : (mg > 1e-6f ? rc : lc);
        }
        ep = P.t_entry[node];
    }
    const float dep = (MC > 0 && P.lmain) ? qrow_l2_mc<MC>(QL, P.lmain + (size_t)ep * P.d, P.d, l) + P.ltail[ep]
                                          : qrow_l2_mc<MC>(QL, P.lifted + (size_t)ep * P.dl, P.dl, l);
    uint32_t nd = 1, rcnt = 0, bsize = 1, nx = 0;
    int buf = 0, ovf = 0;
    if (lane == 0) {
        if constexpr (GV) atomicOr(S.gvis + (ep >> 5), 1u << (ep & 31u));
        else hset_first(S.hash, P.hmask, ep);
        if constexpr (NCH == 0) {
            S.bd0[0] = dep;
            S.bp0[0] = ep;
        }
        if (dep <= S.r2) fifo_put(P, S, 0, ep);
    }
    if (dep <= S.r2) rcnt = 1;
    __syncthreads();
    // (1) beam descent
    if constexpr (NCH > 0) {
        RKeys<NCH> B;
        rk_clear(B);
        if (lane == 0) B.k[0] = pkey(dep, ep);
        auto take_next = [&](bool mark) -> uint32_t { return rk_take(B, mark); };
        // The adjacency of the entry that is next in line once c is marked
        // is fetched while c is expanded; it is used when no new neighbour
        // of c overtakes it.
        uint32_t pc = U32_NONE, pdg = 0, pnb = U32_NONE;
        for (;;) {
            const uint32_t c = take_next(true);
            if (c == U32_NONE) break;
            uint32_t dg, nbr;
            if (c == pc) {
                dg = pdg;
                nbr = pnb;
            } else {
                adj_load<PK>(P, c, dg, nbr);
            }
            pc = take_next(false);
            if (pc != U32_NONE) adj_load<PK>(P, pc, pdg, pnb);
            nx++;
            if (!expand_rb<NCH, MC, PK, BR>(P, S, QL, dg, nbr, rcnt, nd, B, bsize, L, ovf)) break;
        }
    } else for (;;) {
        uint32_t cur = U32_NONE;
        for (uint32_t b = 0; b < bsize; b += 64) {
            uint32_t i = b + lane;
            uint64_t m = ballot(i < bsize && !(S.bp(buf)[i] >> 31));
            if (m) { cur = b + (uint32_t)(__ffsll((long long)m) - 1); break; }
        }
        if (cur == U32_NONE) break;
        uint32_t c = S.bp(buf)[cur];
        __syncthreads();
        if (lane == 0) S.bp(buf)[cur] = c | 0x80000000u;
        __syncthreads();
        nx++;
        if (!expand<true, MC, GV, PK, BR>(P, S, QL, c, rcnt, nd, buf, bsize, L, ovf)) break;
    }
    // (2) flood over collected nodes, FIFO order
    uint32_t head = 0;
    while (!ovf && head < rcnt) {
        uint32_t nb = min((uint32_t)FL, rcnt - head);
        if (!flood_step<MC, FL, GV, PK, BR>(P, S, QL, head, nb, rcnt, nd, ovf)) break;
        head += nb;
        nx += nb;
    }
    if (ovf) {
        if (lane == 0) {
            P.status[qi] = (uint32_t)ovf; P.cand_cnt[qi] = rcnt; P.ndist[qi] = nd; P.pred_cnt[qi] = 0;
            if (P.ovf_count) atomicAdd(P.ovf_count, 1u);
        }
        return;
    }
    // (3) postfilter in base space
    QRow QB;
    qrow_load_mc<MC>(QB, q, P.d, l);
    float qn = qrow_ip_mc<MC>(QB, q, P.d, l);
    uint32_t npred = 0;
    uint32_t* out = P.cand_ids + (size_t)qi * P.out_cap;
    uint8_t* acc = P.accept + (size_t)qi * P.out_cap;
    if constexpr (MC == 0) {
        for (uint32_t b = 0; b < rcnt; b += 4) {
            uint32_t j = b + (uint32_t)(lane >> 4);
            int ok = 0;
            if (j < rcnt) {
                uint32_t o = fifo_get(P, S, j);
                float ip = qrow_ip(QB, P.base + (size_t)o * P.d, P.d, l);
                float key = P.ip ? -ip : fmaf(-2.0f, ip, P.bnorm[o] + qn);
                ok = key <= thr[o];
                if (l == 0 && j < P.out_cap) { out[j] = o; acc[j] = (uint8_t)ok; }
                if (GV && l == 0) S.res[j] = o | (ok ? 0x80000000u : 0u);
            }
            npred += (uint32_t)popc64(ballot(ok && l == 0));
        }
    } else {
        // four candidates per group per pass, loads issued before the reductions
        constexpr int M1 = MC;
        const int g = lane >> 4;
        const int rb = (int)P.d - 16 * M1;
        for (uint32_t b = 0; b < rcnt; b += 16) {
            float x[4][M1], xt[4], bn[4], th[4];
            uint32_t oid[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t s = b + (uint32_t)(g + 4 * j);
                oid[j] = U32_NONE;
                if (s < rcnt) {
                    const uint32_t o = fifo_get(P, S, s);
                    oid[j] = o;
                    if constexpr (BR) {
                        uint32_t w[2];
                        brow_fetch<M1>(P.brows, o, l, w);
#pragma unroll
                        for (int t = 0; t < M1; t++) x[j][t] = (float)brow_byte(w, t);
                        xt[j] = 0.0f;   // d = 16 MC
                    } else {
                        const float* row = P.base + (size_t)o * P.d;
#pragma unroll
                        for (int t = 0; t < M1; t++) x[j][t] = row[16 * t + l];
                        xt[j] = (l < rb) ? row[16 * M1 + l] : 0.0f;
                    }
                    bn[j] = P.ip ? 0.0f : P.bnorm[o];
                    th[j] = thr[o];
                } else {
#pragma unroll
                    for (int t = 0; t < M1; t++) x[j][t] = 0.0f;
                    xt[j] = 0.0f;
                    bn[j] = 0.0f;
                    th[j] = 0.0f;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                float a = 0.0f;
#pragma unroll
                for (int t = 0; t < M1; t++) a = fmaf(QB.v[t], x[j][t], a);
                float ip = tree16(a);
                if (rb) ip += tree16(QB.tail * xt[j]);
                const uint32_t s = b + (uint32_t)(g + 4 * j);
                int ok = 0;
                if (s < rcnt) {
                    float key = P.ip ? -ip : fmaf(-2.0f, ip, bn[j] + qn);
                    ok = key <= th[j];
                    if (l == 0 && s < P.out_cap) { out[s] = oid[j]; acc[s] = (uint8_t)ok; }
                    if (GV && l == 0) S.res[s] = oid[j] | (ok ? 0x80000000u : 0u);
                }
                npred += (uint32_t)popc64(ballot(ok && l == 0));
            }
        }
    }
    if (lane == 0) {
        P.status[qi] = (rcnt > P.out_cap) ? 3u : 0u;
        P.cand_cnt[qi] = rcnt;
        P.ndist[qi] = nd;          // graph distances (the verifier adds rcnt)
        P.pred_cnt[qi] = npred;
        if (P.nexp) P.nexp[qi] = nx;
    }
}

}  // namespace

template <int MC>
__global__ void __launch_bounds__(64) k_range_search(SearchParams P) {
    extern __shared__ __align__(16) char lds[];
    SearchLds S = carve<false>(lds, P);
    if (blockIdx.x >= P.nq) return;
    uint32_t qi = P.qlist ? P.qlist[blockIdx.x] : blockIdx.x;
    search_one<MC, 0, SR_FLOOD_PLAIN>(P, S, qi);
}

// Global tier: block b runs query qlist[b] with its visited bitmap and FIFO
// in global memory (rows b of gvis and gfifo), so neither can fill: the
// reference's flood is unbounded (QueryVisited, include/index.h:33-52, is an
// n-entry stamp array) and so is this one.  Used for the rare queries that
// outgrow the LDS tiers.
template <int MC>
__global__ void __launch_bounds__(64) k_range_search_global(SearchParams P) {
    extern __shared__ __align__(16) char lds[];
    SearchLds S = carve<true>(lds, P);
    if (blockIdx.x >= P.nq) return;
    S.gvis = P.gvis + (size_t)blockIdx.x * P.gstride;
    S.res = P.gfifo + (size_t)blockIdx.x * P.n;
    const uint32_t qi = P.qlist ? P.qlist[blockIdx.x] : blockIdx.x;
    search_one<MC, 0, SR_FLOOD_PLAIN, true>(P, S, qi);
}

// Status list: indices whose status is in `mask` (one block).
__global__ void __launch_bounds__(1024) k_status_list(const uint32_t* __restrict__ status, uint32_t nq, uint32_t mask,
                                                      uint32_t* __restrict__ list, uint32_t* __restrict__ count) {
    __shared__ uint32_t cnt;
    if (threadIdx.x == 0) cnt = 0u;
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < nq; q += blockDim.x) {
        const uint32_t v = status[q];
        if (v < 32u && ((mask >> v) & 1u)) list[atomicAdd(&cnt, 1u)] = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) *count = cnt;
}

// Persistent form: a fixed grid pulls query indices from a device counter.
// Every wave exits once the counter passes nq.
template <int MC, int NCH, bool PK, bool BR>
__global__ void __launch_bounds__(64) k_range_search_persistent(SearchParams P) {
    extern __shared__ __align__(16) char lds[];
    SearchLds S = carve<false>(lds, P);
    __shared__ uint32_t next;
    const uint32_t nq = P.nq_dev ? *P.nq_dev : P.nq;
    for (;;) {
        if (__lane_id() == 0) next = atomicAdd(P.work_counter, 1u);
        __syncthreads();
        uint32_t t = next;
        __syncthreads();
        if (t >= nq) break;
        uint32_t qi = P.qlist ? P.qlist[t] : t;
        search_one<MC, NCH, SR_FLOOD_MAX, false, PK, BR>(P, S, qi);
        __syncthreads();
    }
}

// Search-tree entries for a batch (persistent form): a 16-lane group per
// query, four queries per wavefront, so the descent's dependent loads of
// many queries overlap; per level the node's children, offset and
// hyperplane are fetched together.  Same descent as search_one's
// (rptrees.h:108-123, ties go left).  Queries beyond nq (or *nq_dev) exit.
// key (optional): the lifted distance from the query to its entry, the
// scheduling key of k_entry_order.
// MC = d / 16 = dl / 16 (the search's chunk count) or 0 (run-time form).
template <int MC>
__global__ void __launch_bounds__(64) k_tree_entry_batch(SearchParams P, uint32_t* __restrict__ out,
                                                         float* __restrict__ key) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    const uint32_t nq = P.nq_dev ? *P.nq_dev : P.nq;
    const uint32_t qi = blockIdx.x * 4u + (uint32_t)g;
    if (qi >= nq) return;                       // whole 16-lane groups
    const float* q = P.queries + (size_t)qi * P.d;
    const uint32_t m = P.dl >> 4, r = P.dl & 15;
    auto lifted_q = [&](uint32_t e) -> float {
        if (e < P.d) return sd_to_float(sd_mul(sd_from_float(q[e]), P.scale_bits));
        return e == P.d ? 1.0f : 0.0f;
    };
    QRow QL;
#pragma unroll
    for (int t = 0; t < QMAX_CHUNKS; t++) QL.v[t] = 0.0f;
    if constexpr (MC > 0) {
        // 16 t + l < 16 MC <= d: the query's elements, all requested at once
        float qv[MC];
#pragma unroll
        for (int t = 0; t < MC; t++) qv[t] = q[16 * t + l];
#pragma unroll
        for (int t = 0; t < MC; t++) QL.v[t] = sd_to_float(sd_mul(sd_from_float(qv[t]), P.scale_bits));
    } else {
        for (uint32_t t = 0; t < m; t++) {
            const float v = lifted_q(16u * t + (uint32_t)l);
#pragma unroll
            for (int u = 0; u < QMAX_CHUNKS; u++)
                if ((uint32_t)u == t) QL.v[u] = v;
        }
    }
    QL.tail = ((uint32_t)l < r) ? lifted_q(16u * m + (uint32_t)l) : 0.0f;
    uint32_t node = P.t_root;
    for (;;) {
        const uint32_t left = P.t_left[node], right = P.t_right[node];
        const float off = P.t_off[node];
        const float ip = qrow_ip_mc<MC>(QL, P.t_hyper + (size_t)node * P.dl, P.dl, l);
        if (left == U32_NONE) break;
        const float mg = ip - off;
        node = (mg < -1e-6f) ? left : (mg > 1e-6f ? right : left);
    }
    const uint32_t ep = P.t_entry[node];
    if (l == 0) out[qi] = ep;
    if (key) {
        const float dep = (P.d % 16 == 0 && P.lmain && P.ltail)
                              ? qrow_l2_mc<MC>(QL, P.lmain + (size_t)ep * P.d, P.d, l) + P.ltail[ep]
                              : qrow_l2_mc<MC>(QL, P.lifted + (size_t)ep * P.dl, P.dl, l);
        if (l == 0) key[qi] = dep;
    }
}

// Work-queue order of a batch: queries by ascending key (the lifted distance
// to the entry), so the queries likely to collect many candidates -- and so
// to run the longest floods -- start first and the batch does not wait on a
// late long query.  A counting sort over ORD_BUCKETS buckets of width
// R2 / 64 (the last bucket takes every key from 4 R2 on); order within a
// bucket is arbitrary.  One workgroup.  Every query's results are its own,
// so the order changes only the timing.
#define ORD_BUCKETS 256
__global__ void __launch_bounds__(1024) k_entry_order(const float* __restrict__ key, const uint32_t* nq_dev,
                                                      uint32_t nq_host, float inv_w, uint32_t* __restrict__ qlist) {
    __shared__ uint32_t cnt[ORD_BUCKETS];
    const uint32_t nq = nq_dev ? *nq_dev : nq_host;
    const uint32_t t = threadIdx.x;
    auto bucket = [&](uint32_t q) -> uint32_t {
        const float b = fminf(key[q] * inv_w, (float)(ORD_BUCKETS - 1));   // NaN -> last bucket
        return b >= 0.0f ? (uint32_t)b : 0u;
    };
    for (uint32_t i = t; i < ORD_BUCKETS; i += blockDim.x) cnt[i] = 0u;
    __syncthreads();
    for (uint32_t q = t; q < nq; q += blockDim.x) atomicAdd(&cnt[bucket(q)], 1u);
    __syncthreads();
    if (t == 0) {
        uint32_t run = 0;
        for (int i = 0; i < ORD_BUCKETS; i++) { const uint32_t c = cnt[i]; cnt[i] = run; run += c; }
    }
    __syncthreads();
    for (uint32_t q = t; q < nq; q += blockDim.x) qlist[atomicAdd(&cnt[bucket(q)], 1u)] = q;
}

hipError_t launch_tree_entry_batch(const SearchParams& P, uint32_t max_nq, uint32_t* out, float* key,
                                   hipStream_t s) {
    if (max_nq == 0) return hipSuccess;
    const uint32_t m = P.dl >> 4;
    const int mc = ((P.d >> 4) == m && (m == 8 || m == 6)) ? (int)m : 0;
    if (mc == 8) hipLaunchKernelGGL(k_tree_entry_batch<8>, dim3((max_nq + 3) / 4), dim3(64), 0, s, P, out, key);
    else if (mc == 6) hipLaunchKernelGGL(k_tree_entry_batch<6>, dim3((max_nq + 3) / 4), dim3(64), 0, s, P, out, key);
    else hipLaunchKernelGGL(k_tree_entry_batch<0>, dim3((max_nq + 3) / 4), dim3(64), 0, s, P, out, key);
    return hipGetLastError();
}

// Inner-product metric: each query's R2 = |s q|^2 + 1 + M*^2 (rknn_query.cpp:
// 2719-2729: v_j = double(q_j) * s, qn2s = fma(v_j, v_j, qn2s) in element
// order, then (qn2s + 1) + M*^2), as R2f = the largest float <= R2 -- the
// range search's double(float D) <= R2 test in FP32.  One thread per query,
// in soft binary64 (the sum is sequential); queries beyond nq (or *nq_dev)
// exit.
__global__ void __launch_bounds__(256) k_query_r2(SearchParams P, uint64_t m2_bits, float* __restrict__ r2) {
    const uint32_t nq = P.nq_dev ? *P.nq_dev : P.nq;
    const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    const float* q = P.queries + (size_t)qi * P.d;
    uint64_t qn2s = 0;
    for (uint32_t j = 0; j < P.d; j++) {
        const uint64_t v = sd_mul(sd_from_float(q[j]), P.scale_bits);
        qn2s = sd_fma(v, v, qn2s);
    }
    const uint64_t R2 = sd_add(sd_add(qn2s, 0x3ff0000000000000ull), m2_bits);   // R2 > 0
    float f = sd_to_float(R2);
    // both positive: the bit patterns order the values
    if (sd_from_float(f) > R2) f = __uint_as_float(__float_as_uint(f) - 1u);
    r2[qi] = f;
}

hipError_t launch_query_r2(const SearchParams& P, uint32_t max_nq, double mstar2, float* r2, hipStream_t s) {
    if (max_nq == 0) return hipSuccess;
    uint64_t mb;
    std::memcpy(&mb, &mstar2, 8);
    hipLaunchKernelGGL(k_query_r2, dim3((max_nq + 255) / 256), dim3(256), 0, s, P, mb, r2);
    return hipGetLastError();
}

hipError_t launch_entry_order(const SearchParams& P, const float* key, uint32_t* qlist, hipStream_t s) {
    const float inv_w = 64.0f / fmaxf(P.R2f, 1e-30f);
    hipLaunchKernelGGL(k_entry_order, dim3(1), dim3(1024), 0, s, key, P.nq_dev, P.nq, inv_w, qlist);
    return hipGetLastError();
}

// Split lifted layout: lmain[i] = the first d coordinates of lifted row i
// (rows of d floats, so with d % 16 == 0 a row starts on a 64-byte boundary
// and covers d / 32 128-byte lines instead of straddling one more), and
// ltail[i] = the reduced term of its two lift coordinates against a query's
// (1, 0): tree16 of the lane products ((1 - h0)^2, (0 - h1)^2, 0, ...), the
// value the whole-row distance adds after its main reduction.  A 16-lane
// group per row.
__global__ void k_lift_split(const float* __restrict__ lifted, uint32_t n, uint32_t d, float* __restrict__ lmain,
                             float* __restrict__ ltail) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    const uint32_t i = blockIdx.x * 4u + (uint32_t)g;
    if (i >= n) return;                          // whole 16-lane groups
    const uint32_t dl = d + 2;
    const float* row = lifted + (size_t)i * dl;
    for (uint32_t e = (uint32_t)l; e < d; e += 16) lmain[(size_t)i * d + e] = row[e];
    const float qt = (l == 0) ? 1.0f : 0.0f;
    const float a = (l < 2) ? qt - row[d + l] : 0.0f;
    const float t = tree16(a * a);
    if (l == 0) ltail[i] = t;
}

hipError_t launch_lift_split(const float* lifted, uint32_t n, uint32_t d, float* lmain, float* ltail,
                             hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lift_split, dim3((n + 3) / 4), dim3(64), 0, s, lifted, n, d, lmain, ltail);
    return hipGetLastError();
}

// Byte rows (see brow_fetch): one thread per element.  *bad counts the
// elements that are not one of the floats 0..255 (-0 included) or whose
// lifted coordinate lmain[i][e] is not brow_lifted(b, scale) bit for bit.
__global__ void k_byte_rows(const float* __restrict__ base, const float* __restrict__ lmain, uint32_t n, uint32_t d,
                            double scale, uint8_t* __restrict__ brows, uint32_t* __restrict__ bad) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)n * d) return;
    const uint32_t i = (uint32_t)(t / d), e = (uint32_t)(t % d);
    const uint32_t mc = d >> 4, l = e & 15u, c = e >> 4;
    const float x = base[t];
    const uint32_t b = (x >= 0.0f && x <= 255.0f) ? (uint32_t)x : 0u;
    const bool ok = __float_as_uint(x) == __float_as_uint((float)b) &&
                    __float_as_uint(lmain[t]) == __float_as_uint(brow_lifted(b, bscale(scale)));
    brows[(size_t)i * d + l * mc + c] = (uint8_t)b;
    if (!ok) atomicAdd(bad, 1u);
}

hipError_t launch_byte_rows(const float* base, const float* lmain, uint32_t n, uint32_t d, double scale,
                            uint8_t* brows, uint32_t* bad, hipStream_t s) {
    const uint64_t el = (uint64_t)n * d;
    if (el == 0) return hipSuccess;
    hipLaunchKernelGGL(k_byte_rows, dim3((unsigned)((el + 255) / 256)), dim3(256), 0, s, base, lmain, n, d, scale,
                       brows, bad);
    return hipGetLastError();
}

// Sentinel adjacency: one thread per slot.
__global__ void k_pack_links(const uint32_t* __restrict__ links, const uint32_t* __restrict__ deg, uint32_t n,
                             uint32_t M, uint32_t* __restrict__ plinks) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)n * M) return;
    const uint32_t i = (uint32_t)(t / M), j = (uint32_t)(t % M);
    plinks[t] = j < deg[i] ? links[t] : U32_NONE;
}

hipError_t launch_pack_links(const uint32_t* links, const uint32_t* deg, uint32_t n, uint32_t M, uint32_t* plinks,
                             hipStream_t s) {
    const uint64_t slots = (uint64_t)n * M;
    if (slots == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_links, dim3((unsigned)((slots + 255) / 256)), dim3(256), 0, s, links, deg, n, M,
                       plinks);
    return hipGetLastError();
}

// Chunk count shared by the lifted rows (d + 2) and the base rows (d).
static int search_chunks(const SearchParams& P) {
    const int m = (int)(P.dl >> 4);
    if ((int)(P.d >> 4) != m) return 0;
    return (m == 8 || m == 6) ? m : 0;
}

// Split rows (and the packed adjacency, which needs them): persistent form
// with whole 16-float chunks only.
static void split_layout(SearchParams& P, bool persistent, int mc) {
    if (!persistent || mc == 0 || (P.d & 15u) || !P.ltail) P.lmain = nullptr;
    if (!P.lmain) {
        P.plinks = nullptr;
        P.brows = nullptr;
    }
}

// Register-beam chunks for beams up to Lmax (0: Lmax > 512, LDS beam).
static int search_nch(uint32_t Lmax) {
    return Lmax <= 64 ? 1 : Lmax <= 128 ? 2 : Lmax <= 256 ? 4 : Lmax <= 512 ? 8 : 0;
}

// P.rbeam must already say which form runs (it changes the beam areas).
size_t search_lds_bytes(const SearchParams& P) {
    auto r = [](size_t b) { return (b + 15) & ~size_t(15); };
    const size_t beam = P.rbeam ? r(8 * (size_t)P.Lmax) : 4 * r(4 * (size_t)P.Lmax) + r(8 * 64);
    const size_t sets = P.gvis ? 0 : r(4 * (size_t)(P.hmask + 1)) + r(4 * (size_t)P.rcap);
    return beam + sets + r(4 * SR_TID) + r(4 * SR_TID);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <int MC, int NCH, bool PK, bool BR>
static hipError_t launch_persistent_pk(const SearchParams& P, uint32_t grid, size_t lds, hipStream_t s) {
    const void* f = (const void*)k_range_search_persistent<MC, NCH, PK, BR>;
    if (hipError_t e = set_lds_limit(f, lds); e != hipSuccess) return e;
    hipLaunchKernelGGL((k_range_search_persistent<MC, NCH, PK, BR>), dim3(grid), dim3(64), lds, s, P);
    return hipSuccess;
}

// The persistent kernel for P's layout: packed adjacency and byte rows only
// with the split layout (MC > 0).
template <int MC, int NCH>
static const void* persistent_kernel(const SearchParams& P) {
    if constexpr (MC > 0) {
        if (P.plinks && P.brows) return (const void*)k_range_search_persistent<MC, NCH, true, true>;
        if (P.plinks) return (const void*)k_range_search_persistent<MC, NCH, true, false>;
        if (P.brows) return (const void*)k_range_search_persistent<MC, NCH, false, true>;
    }
    return (const void*)k_range_search_persistent<MC, NCH, false, false>;
}

template <int MC, int NCH>
static hipError_t launch_persistent_t(const SearchParams& P, uint32_t grid, size_t lds, hipStream_t s) {
    if constexpr (MC > 0) {
        if (P.plinks && P.brows) return launch_persistent_pk<MC, NCH, true, true>(P, grid, lds, s);
        if (P.plinks) return launch_persistent_pk<MC, NCH, true, false>(P, grid, lds, s);
        if (P.brows) return launch_persistent_pk<MC, NCH, false, true>(P, grid, lds, s);
    }
    return launch_persistent_pk<MC, NCH, false, false>(P, grid, lds, s);
}

template <int MC>
static hipError_t launch_search_mc(const SearchParams& P0, bool persistent, uint32_t grid, hipStream_t s) {
    SearchParams P = P0;
    P.rbeam = persistent && search_nch(P.Lmax) > 0;
    split_layout(P, persistent, MC);
    size_t lds = search_lds_bytes(P);
    if (persistent) {
        hipError_t e = hipSuccess;
        switch (search_nch(P.Lmax)) {
            case 1: e = launch_persistent_t<MC, 1>(P, grid, lds, s); break;
            case 2: e = launch_persistent_t<MC, 2>(P, grid, lds, s); break;
            case 4: e = launch_persistent_t<MC, 4>(P, grid, lds, s); break;
            case 8: e = launch_persistent_t<MC, 8>(P, grid, lds, s); break;
            default: e = launch_persistent_t<MC, 0>(P, grid, lds, s); break;
        }
        if (e != hipSuccess) return e;
    } else {
        if (P.nq == 0) return hipSuccess;
        if (hipError_t e = set_lds_limit((const void*)k_range_search<MC>, lds); e != hipSuccess) return e;
        hipLaunchKernelGGL(k_range_search<MC>, dim3(P.nq), dim3(64), lds, s, P);
    }
    return hipGetLastError();
}

template <int MC>
static hipError_t launch_global_mc(const SearchParams& P0, hipStream_t s) {
    SearchParams P = P0;
    P.rbeam = 0;
    P.lmain = nullptr;
    P.plinks = nullptr;
    P.brows = nullptr;
    const size_t lds = search_lds_bytes(P);
    if (hipError_t e = set_lds_limit((const void*)k_range_search_global<MC>, lds); e != hipSuccess) return e;
    hipLaunchKernelGGL(k_range_search_global<MC>, dim3(P.nq), dim3(64), lds, s, P);
    return hipGetLastError();
}

hipError_t launch_range_search_global(const SearchParams& P, hipStream_t s) {
    if (P.nq == 0) return hipSuccess;
    if (!P.gvis || !P.gfifo || (P.gstride & 3u) || (uint64_t)P.gstride * 32u < P.n) return hipErrorInvalidValue;
    switch (search_chunks(P)) {
        case 8: return launch_global_mc<8>(P, s);
        case 6: return launch_global_mc<6>(P, s);
        default: return launch_global_mc<0>(P, s);
    }
}

hipError_t launch_status_list(const uint32_t* status, uint32_t nq, uint32_t mask, uint32_t* list, uint32_t* count,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_status_list, dim3(1), dim3(1024), 0, s, status, nq, mask, list, count);
    return hipGetLastError();
}

hipError_t launch_range_search(const SearchParams& P, bool persistent, uint32_t grid, hipStream_t s) {
    switch (search_chunks(P)) {
        case 8: return launch_search_mc<8>(P, persistent, grid, s);
        case 6: return launch_search_mc<6>(P, persistent, grid, s);
        default: return launch_search_mc<0>(P, persistent, grid, s);
    }
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

template <int MC, int NCH>
static int occupancy_t(const SearchParams& P, size_t lds) {
    int occ = 0;
    const void* f = persistent_kernel<MC, NCH>(P);
    if (set_lds_limit(f, lds) != hipSuccess) return 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, f, 64, lds) != hipSuccess) occ = 1;
    return occ;
}

template <int MC>
static int occupancy_mc(const SearchParams& P, size_t lds) {
    switch (search_nch(P.Lmax)) {
        case 1: return occupancy_t<MC, 1>(P, lds);
        case 2: return occupancy_t<MC, 2>(P, lds);
        case 4: return occupancy_t<MC, 4>(P, lds);
        case 8: return occupancy_t<MC, 8>(P, lds);
        default: return occupancy_t<MC, 0>(P, lds);
    }
}

int search_persistent_occupancy(const SearchParams& P0) {
    SearchParams P = P0;
    P.rbeam = search_nch(P.Lmax) > 0;
    split_layout(P, true, search_chunks(P));
    const size_t lds = search_lds_bytes(P);
    switch (search_chunks(P)) {
        case 8: return occupancy_mc<8>(P, lds);
        case 6: return occupancy_mc<6>(P, lds);
        default: return occupancy_mc<0>(P, lds);
    }
}
