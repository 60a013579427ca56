// edge_policy.hpp -- the edge policies of prune_pool_to_ids
// (include/index.h:283-829) for the graph prunes: residual cover, the ball
// ("bubble") order and occlusion in pure and hybrid forms, the lsg
// (locally scaled) metric, and lsg + ball.  One wavefront per row, wave-
// uniform control flow; the double-precision keys and tests are evaluated
// as GCC builds the reference (fused where it fuses them).
#pragma once
#include "common.hpp"
#include "kernels.hpp"

__host__ __device__ __forceinline__ size_t pol_r16(size_t b) { return (b + 15) & ~size_t(15); }
__host__ __device__ __forceinline__ uint32_t pol_p2(uint32_t v) {
    uint32_t p = 64;
    while (p < v) p <<= 1;
    return p;
}

// The bubble optimizer's sizes for pools of up to pmax entries: universe
// (the selected ids and the alternatives), targets, signature words.
struct BubbleDims {
    uint32_t nu, nt, nt2, words;
};
__host__ __device__ __forceinline__ BubbleDims bubble_dims(const EdgePolicyDev& E, uint32_t pmax, uint32_t M) {
    BubbleDims b;
    b.nu = M + (E.bubble_full ? pmax : (E.auto_candidates < pmax ? E.auto_candidates : pmax));
    b.nt = E.bubble_full ? pmax : b.nu;
    b.nt2 = pol_p2(b.nt);
    b.words = (b.nt + 63) / 64;
    return b;
}
__host__ __device__ __forceinline__ size_t bubble_scratch_bytes(const EdgePolicyDev& E, uint32_t pmax, uint32_t M) {
    if (!E.bubble_auto) return 0;
    const BubbleDims b = bubble_dims(E, pmax, M);
    return pol_r16(4 * (size_t)b.nu) + pol_r16(4 * (size_t)b.nt2) + pol_r16(4 * (size_t)b.nt2) +
           pol_r16(8 * (size_t)b.nt2) + pol_r16(8 * (size_t)b.nu) + pol_r16(16 * (size_t)b.nu * b.words) +
           pol_r16(8 * (size_t)b.nt) + pol_r16(4 * (size_t)M) + pol_r16(b.nu) + 3 * pol_r16(8 * (size_t)b.words);
}

// scratch bytes of a policy prune over pools of up to pmax entries (cap2 the
// sorted size): two sorted orders of (u64 key, u32 id, f32 d), and the
// bubble optimizer's arrays
__host__ __device__ __forceinline__ size_t policy_scratch_bytes(uint32_t cap2, const EdgePolicyDev& E, uint32_t M) {
    return pol_r16((size_t)cap2 * (2 * 8 + 2 * 4 + 4)) + bubble_scratch_bytes(E, cap2, M);
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

// optimize_verification_constrained_bubble (:521-723) over the selected ids
// S (the plain prune's): universe = S + the first alternatives of the ball
// order (bid, w of them) not in it; targets = the pool, or the universe's own
// pool entries, by (d, id); greedy swaps (at most min(max_swaps,
// alternatives)) that keep every target some selected edge verifies
// verified and maximise (route coverage gain, power key gain), the first
// such (candidate, slot) pair in scan order.  S's ids are replaced in place.
__device__ void bubble(const float* x, const float* nrm, int d, uint32_t M, float alpha, const EdgePolicyDev& E,
                       uint32_t base, const float* pd, const uint32_t* pi, uint32_t w, const uint32_t* bid,
                       char* scr, uint32_t pmax, Sel& S) {
    const uint32_t lane = (uint32_t)__lane_id();
    const int l = (int)(lane & 15), g = (int)(lane >> 4);
    if (!E.bubble_auto || S.deg == 0 || w == 0) return;
    const BubbleDims B = bubble_dims(E, pmax, M);
    size_t o = 0;
    auto take = [&](size_t bytes) { char* p = scr + o; o += pol_r16(bytes); return p; };
    uint32_t* U = (uint32_t*)take(4 * (size_t)B.nu);
    uint32_t* tid = (uint32_t*)take(4 * (size_t)B.nt2);
    float* td = (float*)take(4 * (size_t)B.nt2);
    uint64_t* tk = (uint64_t*)take(8 * (size_t)B.nt2);
    double* pk = (double*)take(8 * (size_t)B.nu);
    uint64_t* ver = (uint64_t*)take(16 * (size_t)B.nu * B.words);
    uint64_t* rou = ver + (size_t)B.nu * B.words;
    int* vc = (int*)take(8 * (size_t)B.nt);
    int* rc = vc + B.nt;
    uint32_t* sel = (uint32_t*)take(4 * (size_t)M);
    uint8_t* is_sel = (uint8_t*)take(B.nu);
    uint64_t* unc = (uint64_t*)take(8 * (size_t)B.words);
    uint64_t* rc1 = (uint64_t*)take(8 * (size_t)B.words);
    uint64_t* need = (uint64_t*)take(8 * (size_t)B.words);   // required, verified by exactly one
    // universe
    uint32_t nu = S.deg;
    for (uint32_t t = lane; t < S.deg; t += 64) U[t] = S.id[t];
    __syncthreads();
    const uint32_t limit = E.bubble_full ? w : E.auto_candidates;
    uint32_t alts = 0;
    for (uint32_t t = 0; t < w && alts < limit; t++) {
        const uint32_t c = bid[t];
        int in = 0;
        for (uint32_t u0 = 0; u0 < nu && !in; u0 += 64) in = __any(u0 + lane < nu && U[u0 + lane] == c);
        if (in) continue;
        if (lane == 0) U[nu] = c;
        nu++;
        alts++;
        __syncthreads();
    }
    if (alts == 0) return;
    // targets
    uint32_t nt = 0;
    if (E.bubble_full) {
        for (uint32_t t = lane; t < w; t += 64) { tid[t] = pi[t]; td[t] = pd[t]; }
        nt = w;
    } else {
        for (uint32_t u = lane; u < nu; u += 64) {
            float dv = 0.0f;
            for (uint32_t t = 0; t < w; t++)
                if (pi[t] == U[u]) { dv = pd[t]; break; }
            tid[u] = U[u];
            td[u] = dv;
            tk[u] = dkey((double)dv);
        }
        nt = nu;
        __syncthreads();
        sort_keys(tk, tid, td, nt, pol_p2(nt));
    }
    __syncthreads();
    const uint32_t words = (nt + 63) / 64;
    // signatures: verify / route bits of every (universe, target) pair; power keys
    for (uint32_t u = 0; u < nu; u++) {
        const uint32_t id = U[u];
        QRow QU;
        qrow_load(QU, x + (size_t)id * d, d, l);
        const float nu_ = nrm[id];
        if (lane == 0) {
            float db = 0.0f;
            for (uint32_t t = 0; t < w; t++)
                if (pi[t] == id) { db = pd[t]; break; }
            pk[u] = fma(-E.lambda, sort_radj(E, id), (double)db);
        }
        // 64 targets per word: each 16-lane group takes every fourth one,
        // the groups' bits are merged in registers and one lane stores the
        // word (plain stores, as every later read of them)
        for (uint32_t t0 = 0; t0 < nt; t0 += 64) {
            uint64_t vb = 0, rb = 0;
            for (uint32_t s4 = 0; s4 < 64; s4 += 4) {
                const uint32_t t = t0 + s4 + (uint32_t)g;
                if (t < nt) {
                    const uint32_t tg = tid[t];
                    float dsc = 0.0f;
                    if (tg != id) {
                        const float ip = qrow_ip(QU, x + (size_t)tg * d, d, l);
                        dsc = d2_combine(ip, nu_, nrm[tg]);
                    }
                    if (tg == id || alpha * dsc < td[t]) vb |= 1ull << (s4 + (uint32_t)g);
                    if (ball_occ(E, alpha, base, id, td[t], dsc)) rb |= 1ull << (s4 + (uint32_t)g);
                }
            }
            vb |= __shfl_xor(vb, 16);
            vb |= __shfl_xor(vb, 32);
            rb |= __shfl_xor(rb, 16);
            rb |= __shfl_xor(rb, 32);
            if (lane == 0) {
                ver[(size_t)u * B.words + (t0 >> 6)] = vb;
                rou[(size_t)u * B.words + (t0 >> 6)] = rb;
            }
        }
        __syncthreads();
    }
    // the selected edges, in their order
    for (uint32_t u = lane; u < nu; u += 64) is_sel[u] = 0;
    __syncthreads();
    uint32_t ns = 0;
    for (uint32_t s = 0; s < S.deg; s++) {
        const uint32_t id = S.id[s];
        uint32_t found = U32_NONE;
        for (uint32_t u0 = 0; u0 < nu && found == U32_NONE; u0 += 64) {
            const uint64_t m = __ballot(u0 + lane < nu && U[u0 + lane] == id);
            if (m) found = u0 + (uint32_t)(__ffsll((long long)m) - 1);
        }
        if (found != U32_NONE) {
            if (lane == 0) { sel[ns] = found; is_sel[found] = 1; }
            ns++;
        }
        __syncthreads();
    }
    if (ns == 0) return;
    auto bit = [&](const uint64_t* b, uint32_t u, uint32_t t) -> int {
        return (int)((b[(size_t)u * B.words + (t >> 6)] >> (t & 63u)) & 1ull);
    };
    for (uint32_t t = lane; t < nt; t += 64) {
        int a = 0, r = 0;
        for (uint32_t s = 0; s < ns; s++) { a += bit(ver, sel[s], t); r += bit(rou, sel[s], t); }
        vc[t] = a;
        rc[t] = r;
    }
    __syncthreads();
    // required targets: verified before any swap (fixed from here on)
    uint64_t* req = tk;   // the target keys are no longer needed
    for (uint32_t wd = lane; wd < words; wd += 64) {
        uint64_t m = 0;
        for (uint32_t b2 = 0; b2 < 64 && wd * 64 + b2 < nt; b2++)
            if (vc[wd * 64 + b2] > 0) m |= 1ull << b2;
        req[wd] = m;
    }
    __syncthreads();
    const uint32_t max_swaps = min(E.auto_max_swaps, alts);
    for (uint32_t swaps = 0; swaps < max_swaps; swaps++) {
        for (uint32_t wd = lane; wd < words; wd += 64) {
            uint64_t a = 0, b1 = 0, c1 = 0;
            for (uint32_t b2 = 0; b2 < 64 && wd * 64 + b2 < nt; b2++) {
                const uint32_t t = wd * 64 + b2;
                if (rc[t] == 0) a |= 1ull << b2;
                if (rc[t] == 1) b1 |= 1ull << b2;
                if (vc[t] == 1) c1 |= 1ull << b2;
            }
            unc[wd] = a;
            rc1[wd] = b1;
            need[wd] = c1 & req[wd];
        }
        __syncthreads();
        // every (candidate q, slot) pair in scan order, one per lane at a time
        long best_rg = LONG_MIN;
        double best_pg = -INFINITY;
        uint32_t best_p = U32_NONE;
        const uint32_t npairs = nu * ns;
        for (uint32_t p = lane; p < npairs; p += 64) {
            const uint32_t q = p / ns, s = p % ns;
            if (is_sel[q]) continue;
            const uint32_t old = sel[s];
            bool feasible = true;
            long newly = 0, lost = 0;
            for (uint32_t wd = 0; wd < words; wd++) {
                const uint64_t ov = ver[(size_t)old * B.words + wd], qv = ver[(size_t)q * B.words + wd];
                const uint64_t orr = rou[(size_t)old * B.words + wd], qr = rou[(size_t)q * B.words + wd];
                if (need[wd] & ov & ~qv) { feasible = false; break; }
                newly += __popcll(qr & unc[wd]);
                lost += __popcll(orr & rc1[wd] & ~qr);
            }
            if (!feasible) continue;
            const long rg = newly - lost;
            const double pg = pk[old] - pk[q];
            if (rg > best_rg || (rg == best_rg && pg > best_pg)) { best_rg = rg; best_pg = pg; best_p = p; }
        }
        // the wave's best: largest (gain, power gain), then the first pair
        for (int off = 32; off > 0; off >>= 1) {
            const long org = __shfl_xor(best_rg, off);
            const double opg = __shfl_xor(best_pg, off);
            const uint32_t op = __shfl_xor(best_p, off);
            const bool take_o = op != U32_NONE &&
                                (best_p == U32_NONE || org > best_rg || (org == best_rg && opg > best_pg) ||
                                 (org == best_rg && opg == best_pg && op < best_p));
            if (take_o) { best_rg = org; best_pg = opg; best_p = op; }
        }
        if (best_p == U32_NONE || best_rg < 0 || (best_rg == 0 && best_pg <= 1e-12)) break;
        const uint32_t bq = best_p / ns, bs = best_p % ns, old = sel[bs];
        for (uint32_t t = lane; t < nt; t += 64) {
            vc[t] += bit(ver, bq, t) - bit(ver, old, t);
            rc[t] += bit(rou, bq, t) - bit(rou, old, t);
        }
        __syncthreads();
        if (lane == 0) { is_sel[old] = 0; is_sel[bq] = 1; sel[bs] = bq; }
        __syncthreads();
    }
    for (uint32_t s = lane; s < ns; s += 64) { S.id[s] = U[sel[s]]; S.out[s] = U[sel[s]]; }
    __syncthreads();
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
    char* bscr = scratch + pol_r16((size_t)cap2 * (2 * 8 + 2 * 4 + 4));
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
    if (E.bubble_auto) {
        build_ball();
        extend<0>(x, nrm, d, alpha, E, base, pi, pd, nullptr, w, M, S);
        bubble(x, nrm, d, M, alpha, E, base, pd, pi, w, bid, bscr, cap2, S);
    } else if (E.lsg && E.ball && E.lsg_alpha != 0.0) {
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
