// edge_policy.hpp -- the edge policies of prune_pool_to_ids
// (include/index.h:283-829) for the graph prunes: residual cover, the ball
// ("bubble") order and occlusion in pure and hybrid forms, the lsg
// (locally scaled) metric, and lsg + ball.  One wavefront per row, wave-
// uniform control flow; the double-precision keys and tests are evaluated
// as GCC builds the reference (fused where it fuses them).
#pragma once
#include "common.hpp"
#include "kernels.hpp"

// scratch bytes per pool entry: two sorted orders of (u64 key, u32 id, f32 d)
__host__ __device__ __forceinline__ size_t policy_scratch_bytes(uint32_t cap2) {
    return (size_t)cap2 * (2 * 8 + 2 * 4 + 4);
}

namespace edgepol {

// doubles as ordered 64-bit keys (-0 taken as +0, as the comparison does)
__device__ __forceinline__ uint64_t dkey(double v) {
    if (v == 0.0) v = 0.0;
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dval(uint64_t k) {
    const uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// edge_radj (:291-293)
__device__ __forceinline__ double radj(const EdgePolicyDev& E, uint32_t id, int k) {
    return E.scale2 * (double)E.resid[(size_t)id * E.nk + (size_t)(k - 1)];
}
// edge_sort_radj (:295-303)
__device__ __forceinline__ double sort_radj(const EdgePolicyDev& E, uint32_t id) {
    if (!E.ball || E.nks == 0) return 0.0;
    double best = -INFINITY;
    for (uint32_t t = 0; t < E.nks; t++) best = fmax(best, radj(E, id, E.ks[t]));
    return isfinite(best) ? best : 0.0;
}
// ball_occludes_distance (:474-490): both sides fused
__device__ __forceinline__ int ball_occ(const EdgePolicyDev& E, float alpha, uint32_t base, uint32_t sid, float dbase,
                                        float dsc) {
    if (!E.ball || E.nks == 0) return 0;
    int any = 0, all = 1;
    for (uint32_t t = 0; t < E.nks; t++) {
        const double lhs = fma(-E.lambda, radj(E, sid, E.ks[t]), (double)alpha * (double)dsc);
        const double rhs = fma(-E.lambda, radj(E, base, E.ks[t]), (double)dbase);
        const int occ = lhs < rhs;
        any = any || occ;
        all = all && occ;
        if (!E.occlude_all && any) return 1;
        if (E.occlude_all && !all) return 0;
    }
    return E.occlude_all ? all : any;
}
// lsg_distance2 (:832-838); a weight power of 1 is the product itself
__device__ __forceinline__ double lsg(const EdgePolicyDev& E, uint32_t a, uint32_t b, double raw) {
    if (!E.lsg || E.lsg_alpha == 0.0) return raw;
    const double mp = (double)E.mu[a] * (double)E.mu[b];
    const double den = E.lsg_alpha == 1.0 ? mp : pow(mp, E.lsg_alpha);
    if (!isfinite(den) || den <= 0.0) return raw;
    return raw / den;
}

// Bitonic sort of (k, id, dv) [0, m) by (k, id) in place; m <= cap2 (power of 2).
__device__ void sort_keys(uint64_t* k, uint32_t* id, float* dv, uint32_t m, uint32_t cap2) {
    const uint32_t lane = (uint32_t)__lane_id();
    for (uint32_t t = m + lane; t < cap2; t += 64) { k[t] = ~0ull; id[t] = U32_NONE; if (dv) dv[t] = 0.0f; }
    __syncthreads();
    for (uint32_t s = 2; s <= cap2; s <<= 1) {
        for (uint32_t j = s >> 1; j > 0; j >>= 1) {
            for (uint32_t t = lane; t < cap2; t += 64) {
                const uint32_t p = t ^ j;
                if (p > t) {
                    const bool gt = k[t] > k[p] || (k[t] == k[p] && id[t] > id[p]);
                    if (gt == ((t & s) == 0)) {
                        const uint64_t fk = k[t]; k[t] = k[p]; k[p] = fk;
                        const uint32_t fi = id[t]; id[t] = id[p]; id[p] = fi;
                        if (dv) { const float fd = dv[t]; dv[t] = dv[p]; dv[p] = fd; }
                    }
                }
            }
            __syncthreads();
        }
    }
}

// The selected set with its vectors staged (as rng_select does).
struct Sel {
    float* v;          // M x d
    uint32_t* id;      // M
    float* nrm;        // M
    uint32_t* out;     // the row's output ids
    uint32_t deg;
};

__device__ __forceinline__ int has(const Sel& S, uint32_t cid) {
    const uint32_t lane = (uint32_t)__lane_id();
    for (uint32_t t0 = 0; t0 < S.deg; t0 += 64)
        if (__any(t0 + lane < S.deg && S.id[t0 + lane] == cid)) return 1;
    return 0;
}

__device__ __forceinline__ void append(Sel& S, const QRow& QC, uint32_t cid, float nc, int d) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    if (g == 0) {
        const int mm = d >> 4, r = d & 15;
#pragma unroll
        for (int c = 0; c < QMAX_CHUNKS; c++)
            if (c < mm) S.v[(size_t)S.deg * d + 16 * c + l] = QC.v[c];
        if (l < r) S.v[(size_t)S.deg * d + 16 * mm + l] = QC.tail;
    }
    if (lane == 0) { S.id[S.deg] = cid; S.nrm[S.deg] = nc; S.out[S.deg] = cid; }
    S.deg++;
    __syncthreads();
}

// One extension pass over an order: KIND 0 rng_extend (:267-281, float
// test alpha * d2(s, c) < d2(base, c)), 1 rng_extend_ball (:505-519), 2
// rng_extend_lsg (:853-871, alpha * lsg(d2(s, c)) < the order's key).
template <int KIND>
__device__ void extend(const float* x, const float* nrm, int d, float alpha, const EdgePolicyDev& E, uint32_t base,
                       const uint32_t* ids, const float* dv, const uint64_t* keys, uint32_t m, uint32_t cap, Sel& S) {
    const int lane = __lane_id(), l = lane & 15, g = lane >> 4;
    for (uint32_t t = 0; t < m; t++) {
        if (S.deg >= cap) break;
        const uint32_t cid = ids[t];
        if (has(S, cid)) continue;
        QRow QC;
        qrow_load(QC, x + (size_t)cid * d, d, l);
        const float nc = nrm[cid];
        const float dbase = KIND == 2 ? 0.0f : dv[t];
        const double key = KIND == 2 ? dval(keys[t]) : 0.0;
        int occ = 0;
        for (uint32_t s0 = 0; s0 < S.deg; s0 += 4) {
            const uint32_t s = s0 + (uint32_t)g;
            int o = 0;
            if (s < S.deg) {
                const float ip = qrow_ip(QC, S.v + (size_t)s * d, d, l);
                const float dsc = d2_combine(ip, S.nrm[s], nc);
                if (KIND == 0) o = alpha * dsc < dbase;
                else if (KIND == 1) o = ball_occ(E, alpha, base, S.id[s], dbase, dsc);
                else o = (double)alpha * lsg(E, S.id[s], cid, (double)dsc) < key;
            }
            if (__any(o)) { occ = 1; break; }
        }
        if (!occ) append(S, QC, cid, nc, d);
    }
}

// add_residual_cover_edges (:441-472): per rank, the unselected pool entries
// of positive residual ordered by d - lambda scale^2 residual (fused), up
// to the per-rank share of the budget, without occlusion.
__device__ void cover(const float* x, const float* nrm, uint32_t n, int d, uint32_t M, const EdgePolicyDev& E,
                      uint32_t base, const float* pd, const uint32_t* pi, uint32_t w, uint64_t* sk, uint32_t* sid,
                      uint32_t cap2, Sel& S) {
    const uint32_t lane = (uint32_t)__lane_id();
    const int l = (int)(lane & 15);
    if (!E.resid_cover || S.deg >= M) return;
    const uint32_t total = min(E.budget, M - S.deg);
    if (total == 0 || E.nks == 0) return;
    const uint32_t per_k = max(1u, (total + E.nks - 1) / E.nks);
    uint32_t added = 0;
    for (uint32_t kt = 0; kt < E.nks; kt++) {
        if (added >= total || S.deg >= M) break;
        const int k = E.ks[kt];
        uint32_t c = 0;
        for (uint32_t t0 = 0; t0 < w; t0 += 64) {
            const uint32_t t = t0 + lane;
            const uint32_t cid = t < w ? pi[t] : U32_NONE;
            int ok = cid != U32_NONE && cid < n && cid != base;
            if (ok) {
                for (uint32_t s = 0; s < S.deg; s++)
                    if (S.id[s] == cid) { ok = 0; break; }
            }
            double score = 0.0;
            if (ok) {
                const double r = (double)E.resid[(size_t)cid * E.nk + (size_t)(k - 1)];
                const double risk = r > 0.0 ? r : 0.0;
                ok = risk > 0.0;
                score = fma(-(E.lambda * E.scale2), risk, (double)pd[t]);
            }
            const uint64_t bm = __ballot(ok);
            __syncthreads();
            if (ok) { sk[c + lanes_below(bm)] = dkey(score); sid[c + lanes_below(bm)] = cid; }
            c += (uint32_t)popc64(bm);
            __syncthreads();
        }
        uint32_t c2 = 64;
        while (c2 < c) c2 <<= 1;
        sort_keys(sk, sid, nullptr, c, c2);
        uint32_t ak = 0;
        for (uint32_t u = 0; u < c; u++) {
            if (S.deg >= M || added >= total || ak >= per_k) break;
            const uint32_t cid = sid[u];
            if (has(S, cid)) continue;
            QRow QC;
            qrow_load(QC, x + (size_t)cid * d, d, l);
            append(S, QC, cid, nrm[cid], d);
            ak++;
            added++;
        }
    }
    (void)cap2;
}

}  // namespace edgepol

// prune_pool_to_ids (:772-829) under an edge policy, over a (dist, id)-sorted
// duplicate-free pool of w entries.  scratch: policy_scratch_bytes(cap2),
// cap2 >= max(64, w) a power of two.  Returns the degree.
__device__ uint32_t policy_select(const float* x, const float* nrm, uint32_t n, int d, uint32_t M, float alpha,
                                  const EdgePolicyDev& E, uint32_t base, const float* pd, const uint32_t* pi,
                                  uint32_t w, char* scratch, uint32_t cap2, float* selv, uint32_t* sel,
                                  float* selnrm, uint32_t* out) {
    using namespace edgepol;
    const uint32_t lane = (uint32_t)__lane_id();
    uint64_t* lk = (uint64_t*)scratch;
    uint64_t* bk = lk + cap2;
    uint32_t* lid = (uint32_t*)(bk + cap2);
    uint32_t* bid = lid + cap2;
    float* bd = (float*)(bid + cap2);
    Sel S{selv, sel, selnrm, out, 0};
    const uint32_t bud = min(E.budget, M);
    auto build_lsg = [&]() {
        for (uint32_t t = lane; t < w; t += 64) { lk[t] = dkey(lsg(E, base, pi[t], (double)pd[t])); lid[t] = pi[t]; }
        sort_keys(lk, lid, nullptr, w, cap2);
    };
    auto build_ball = [&]() {
        for (uint32_t t = lane; t < w; t += 64) {
            bk[t] = dkey(fma(-E.lambda, sort_radj(E, pi[t]), (double)pd[t]));
            bid[t] = pi[t];
            bd[t] = pd[t];
        }
        sort_keys(bk, bid, bd, w, cap2);
    };
    if (E.lsg && E.ball && E.lsg_alpha != 0.0) {
        build_lsg();
        build_ball();
        const uint32_t cap = (M - bud) ? (M - bud) : M;
        extend<2>(x, nrm, d, alpha, E, base, lid, nullptr, lk, w, cap, S);
        extend<1>(x, nrm, d, alpha, E, base, bid, bd, nullptr, w, M, S);
        extend<2>(x, nrm, d, alpha, E, base, lid, nullptr, lk, w, M, S);
    } else if (E.lsg && E.lsg_alpha != 0.0) {
        build_lsg();
        extend<2>(x, nrm, d, alpha, E, base, lid, nullptr, lk, w, M, S);
    } else if (E.ball) {
        build_ball();
        if (E.ball_pure) {
            extend<1>(x, nrm, d, alpha, E, base, bid, bd, nullptr, w, M, S);
            extend<0>(x, nrm, d, alpha, E, base, pi, pd, nullptr, w, M, S);
        } else {
            const uint32_t cap = (M - bud) ? (M - bud) : M;
            extend<0>(x, nrm, d, alpha, E, base, pi, pd, nullptr, w, cap, S);
            extend<1>(x, nrm, d, alpha, E, base, bid, bd, nullptr, w, M, S);
            extend<0>(x, nrm, d, alpha, E, base, pi, pd, nullptr, w, M, S);
        }
    } else {
        const uint32_t cb = E.resid_cover ? bud : 0;
        const uint32_t cap = (M - cb) ? (M - cb) : M;
        extend<0>(x, nrm, d, alpha, E, base, pi, pd, nullptr, w, cap, S);
        cover(x, nrm, n, d, M, E, base, pd, pi, w, lk, lid, cap2, S);
        extend<0>(x, nrm, d, alpha, E, base, pi, pd, nullptr, w, M, S);
    }
    return S.deg;
}
