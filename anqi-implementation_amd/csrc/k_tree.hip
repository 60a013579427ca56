// k_tree.hip -- the search tree over the lifted vectors, drawn exactly as the
// reference draws it (include/rptrees.h:28-163 with include/utils.h:169-173).
//
// The reference builds this one tree recursively, depth first, and every
// random choice comes from one std::mt19937 seeded once (a thread_local engine
// seeded by its first call), through libstdc++'s uniform_int_distribution,
// which for a 32-bit engine is Lemire's nearly-divisionless method.  A node's
// draws therefore sit at a stream position fixed by every node before it in
// depth-first pre-order, so the build is a chain over nodes.  One workgroup
// of 1024 threads walks that chain: each node's margins, tie draws and stable
// partition are spread over the workgroup; the engine's output stream is
// generated on the host (it does not depend on the data) and read here.
//
// Per internal node (m points, d dims): draws r0, r1 (rptrees.h:32-34);
// hyperplane h = a - b and midpoint (a + b) / 2 (:56-60); offset = <h, mid>
// (:61); margin = <h, x> - offset (:70-71) with the reference's FP32 inner
// product; margin < -EPS left, > EPS right, else one draw u >> 31 in index
// order (a range-2 Lemire draw never rejects) (:73-75); all on one side ->
// one draw per point for a random deal (:78-83).  Children: left then right,
// left partition order preserved (push_back order).
//
// Node ids are pre-order (root 0); the reference's post-order numbering is
// produced when the tree is written to a file.
#include "common.hpp"
#include "kernels.hpp"

#define TB_THREADS 1024
#define TB_WAVES (TB_THREADS / 64)
#define TB_EPS 1e-6f   // include/distance.h:14

struct TbFrame {
    uint32_t start, m, depth, parent, right;
};

__device__ __forceinline__ bool tb_lemire(const uint32_t* __restrict__ s, uint64_t len, uint64_t& pos,
                                          uint32_t range, uint32_t& out) {
    if (pos >= len) return false;
    uint64_t p = (uint64_t)s[pos++] * range;
    uint32_t low = (uint32_t)p;
    if (low < range) {
        uint32_t t = (0u - range) % range;
        while (low < t) {
            if (pos >= len) return false;
            p = (uint64_t)s[pos++] * range;
            low = (uint32_t)p;
        }
    }
    out = (uint32_t)(p >> 32);
    return true;
}

// Exclusive block-wide prefix of two flags at once; totals returned.
__device__ __forceinline__ void tb_scan2(int fa, int fb, uint32_t* wsum, uint32_t& pa, uint32_t& pb,
                                         uint32_t& ta, uint32_t& tb) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    uint64_t ba = ballot(fa), bb = ballot(fb);
    if (lane == 0) {
        wsum[w] = (uint32_t)popc64(ba);
        wsum[TB_WAVES + w] = (uint32_t)popc64(bb);
    }
    __syncthreads();
    uint32_t sa = 0, sb = 0, xa = 0, xb = 0;
#pragma unroll
    for (int i = 0; i < TB_WAVES; i++) {
        uint32_t ca = wsum[i], cb = wsum[TB_WAVES + i];
        if (i < w) { sa += ca; sb += cb; }
        xa += ca;
        xb += cb;
    }
    __syncthreads();
    pa = sa + (uint32_t)lanes_below(ba);
    pb = sb + (uint32_t)lanes_below(bb);
    ta = xa;
    tb = xb;
}

// side[start + i] for the m points of a node: each 16-lane group takes four
// points per pass (i = g, g + 64, g + 128, g + 192 of the pass) and issues all
// their loads before reducing, so one workgroup keeps many rows in flight.
// Same FMA order and reduction tree as qrow_ip.
template <int MC>
__device__ __forceinline__ void tb_margins(const TreeBuildParams& P, const QRow& hq, float off, uint32_t start,
                                           uint32_t m, uint32_t g, uint32_t l) {
    constexpr uint32_t G = TB_THREADS / 16;
    const uint32_t d = P.d;
    if constexpr (MC == 0) {
        for (uint32_t i = g; i < m; i += G) {
            const float* xr = P.x + (size_t)P.perm[start + i] * d;
            float mg = qrow_ip(hq, xr, (int)d, (int)l) - off;
            if (l == 0) P.side[start + i] = mg < -TB_EPS ? 0 : (mg > TB_EPS ? 1 : 2);
        }
    } else {
        const int r = (int)d - 16 * MC;
        for (uint32_t i0 = 0; i0 < m; i0 += 4 * G) {
            float xv[4][MC], xt[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const uint32_t i = i0 + g + G * (uint32_t)j;
                if (i < m) {
                    const float* xr = P.x + (size_t)P.perm[start + i] * d;
#pragma unroll
                    for (int t = 0; t < MC; t++) xv[j][t] = xr[16 * t + l];
                    xt[j] = ((int)l < r) ? xr[16 * MC + l] : 0.0f;
                } else {
#pragma unroll
                    for (int t = 0; t < MC; t++) xv[j][t] = 0.0f;
                    xt[j] = 0.0f;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                float acc = 0.0f;
#pragma unroll
                for (int t = 0; t < MC; t++) acc = fmaf(hq.v[t], xv[j][t], acc);
                float ip = tree16(acc);
                if (r) ip += tree16(hq.tail * xt[j]);
                const float mg = ip - off;
                const uint32_t i = i0 + g + G * (uint32_t)j;
                if (l == 0 && i < m) P.side[start + i] = mg < -TB_EPS ? 0 : (mg > TB_EPS ? 1 : 2);
            }
        }
    }
}

template <int MC>
__global__ __launch_bounds__(TB_THREADS) void k_tree_build(TreeBuildParams P) {
    extern __shared__ float tb_dyn[];
    float* hyp = tb_dyn;
    float* mid = tb_dyn + P.d;
    __shared__ TbFrame stk[TB_STACK];
    __shared__ uint32_t wsum[2 * TB_WAVES];
    __shared__ TbFrame s_cur;
    __shared__ uint32_t s_top, s_nodes, s_err, s_id, s_what, s_a, s_b;
    __shared__ uint64_t s_pos;
    __shared__ float s_off;
    const uint32_t tid = threadIdx.x, d = P.d;
    const uint32_t g = tid >> 4, l = tid & 15;
    if (tid == 0) {
        stk[0] = TbFrame{0u, P.n, P.max_depth, U32_NONE, 0u};
        s_top = 1;
        s_nodes = 0;
        s_err = 0;
        s_pos = 0;
    }
    __syncthreads();
    for (;;) {
        if (tid == 0) {
            uint32_t what = 2;   // 0 split, 1 leaf, 2 stop
            if (s_top > 0 && s_err == 0) {
                TbFrame f = stk[--s_top];
                s_cur = f;
                uint32_t id = s_nodes;
                if (id >= P.node_cap) {
                    s_err = TB_ERR_NODES;
                } else {
                    s_nodes = id + 1;
                    s_id = id;
                    if (f.parent != U32_NONE) (f.right ? P.right : P.left)[f.parent] = id;
                    if (f.m <= P.leaf || f.depth == 0) {
                        P.left[id] = U32_NONE;
                        P.right[id] = U32_NONE;
                        P.offset[id] = F32_BIG;
                        P.leaf_start[id] = f.start;
                        P.leaf_size[id] = (f.depth == 0 && f.m > P.leaf) ? P.leaf : f.m;
                        what = 1;
                    } else {
                        uint64_t pos = s_pos;
                        uint32_t r0 = 0, r1 = 0;
                        if (!tb_lemire(P.stream, P.stream_len, pos, f.m, r0) ||
                            !tb_lemire(P.stream, P.stream_len, pos, f.m, r1)) {
                            s_err = TB_ERR_STREAM;
                        } else {
                            if (r0 == r1) r0 = (r1 + 1) % f.m;
                            s_a = P.perm[f.start + r0];
                            s_b = P.perm[f.start + r1];
                            s_pos = pos;
                            P.leaf_start[id] = 0;
                            P.leaf_size[id] = 0;
                            what = 0;
                        }
                    }
                }
            }
            if (s_err) what = 2;
            s_what = what;
        }
        __syncthreads();
        const uint32_t what = s_what;
        const TbFrame f = s_cur;
        const uint32_t id = s_id, ia = s_a, ib = s_b;
        uint64_t pos = s_pos;
        __syncthreads();
        if (what == 2) break;
        if (what == 1) continue;

        // hyperplane and midpoint (rptrees.h:56-60)
        const float* xa = P.x + (size_t)ia * d;
        const float* xb = P.x + (size_t)ib * d;
        for (uint32_t j = tid; j < d; j += TB_THREADS) {
            float a = xa[j], b = xb[j];
            float h = a - b;
            mid[j] = (a + b) / 2.0f;
            hyp[j] = h;
            P.hyper[(size_t)id * d + j] = h;
        }
        __syncthreads();
        if (tid < 16) {
            float o = g16_ip(hyp, mid, (int)d, (int)l);
            if (tid == 0) {
                s_off = o;
                P.offset[id] = o;
            }
        }
        __syncthreads();
        const float off = s_off;
        QRow hq;
        qrow_load(hq, hyp, (int)d, (int)l);
        // margins (rptrees.h:64-76); 2 marks a tie to be drawn in index order
        tb_margins<MC>(P, hq, off, f.start, f.m, g, l);
        __syncthreads();
        // resolve ties with consecutive draws; count the left side
        uint32_t nl = 0, ties = 0;
        bool fail = false;
        for (uint32_t base = 0; base < f.m; base += TB_THREADS) {
            uint32_t i = base + tid;
            uint32_t sd = i < f.m ? P.side[f.start + i] : 1u;
            uint32_t tr, lr, tt, tl;
            tb_scan2(sd == 2, sd == 0, wsum, tr, lr, tt, tl);
            if (pos + ties + tt > P.stream_len) { fail = true; break; }
            if (sd == 2) {
                sd = (P.stream[pos + ties + tr] >> 31) ? 1u : 0u;
                P.side[f.start + i] = (uint8_t)sd;
            }
            uint32_t lr2, dummy, tl2, td;
            tb_scan2(i < f.m && sd == 0, 0, wsum, lr2, dummy, tl2, td);
            nl += tl2;
            ties += tt;
        }
        pos += ties;
        if (!fail && (nl == 0 || nl == f.m)) {
            // every point on one side: deal each point by one draw (rptrees.h:78-83)
            if (pos + f.m > P.stream_len) {
                fail = true;
            } else {
                nl = 0;
                for (uint32_t base = 0; base < f.m; base += TB_THREADS) {
                    uint32_t i = base + tid;
                    uint32_t sd = 1;
                    if (i < f.m) {
                        sd = (P.stream[pos + i] >> 31) ? 1u : 0u;
                        P.side[f.start + i] = (uint8_t)sd;
                    }
                    uint32_t lr2, dummy, tl2, td;
                    tb_scan2(i < f.m && sd == 0, 0, wsum, lr2, dummy, tl2, td);
                    nl += tl2;
                }
                pos += f.m;
            }
        }
        if (fail) {
            if (tid == 0) s_err = TB_ERR_STREAM;
            __syncthreads();
            continue;
        }
        // stable partition: left part keeps order, right part follows
        if (f.m <= TB_THREADS) {
            uint32_t i = tid;
            uint32_t v = 0, sd = 1;
            if (i < f.m) {
                v = P.perm[f.start + i];
                sd = P.side[f.start + i];
            }
            uint32_t lr, dummy, tl, td;
            tb_scan2(i < f.m && sd == 0, 0, wsum, lr, dummy, tl, td);
            if (i < f.m) {
                uint32_t dst = sd == 0 ? lr : nl + (i - lr);
                P.perm[f.start + dst] = v;
            }
        } else {
            uint32_t carry = 0;
            for (uint32_t base = 0; base < f.m; base += TB_THREADS) {
                uint32_t i = base + tid;
                uint32_t sd = i < f.m ? P.side[f.start + i] : 1u;
                uint32_t lr, dummy, tl, td;
                tb_scan2(i < f.m && sd == 0, 0, wsum, lr, dummy, tl, td);
                if (i < f.m) {
                    uint32_t left_before = carry + lr;
                    uint32_t dst = sd == 0 ? left_before : nl + (i - left_before);
                    P.tmp[f.start + dst] = P.perm[f.start + i];
                }
                carry += tl;
            }
            __syncthreads();
            for (uint32_t i = tid; i < f.m; i += TB_THREADS) P.perm[f.start + i] = P.tmp[f.start + i];
        }
        __syncthreads();
        if (tid == 0) {
            s_pos = pos;
            if (s_top + 2 > TB_STACK) {
                s_err = TB_ERR_STACK;
            } else {
                stk[s_top++] = TbFrame{f.start + nl, f.m - nl, f.depth - 1, id, 1u};
                stk[s_top++] = TbFrame{f.start, nl, f.depth - 1, id, 0u};
            }
        }
        __syncthreads();
    }
    if (tid == 0) {
        P.status[0] = s_nodes;
        P.status[1] = s_err;
        P.status[2] = (uint32_t)s_pos;
        P.status[3] = (uint32_t)(s_pos >> 32);
    }
}

__global__ void k_tree_entries(const uint32_t* __restrict__ left, const uint32_t* __restrict__ leaf_start,
                               const uint32_t* __restrict__ leaf_size, const uint32_t* __restrict__ perm,
                               uint32_t nodes, uint32_t* __restrict__ entry) {
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nodes) return;
    entry[k] = (left[k] == U32_NONE && leaf_size[k]) ? perm[leaf_start[k]] : 0u;
}

__global__ void k_iota_u32(uint32_t* __restrict__ p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

hipError_t launch_tree_build(const TreeBuildParams& P, hipStream_t s) {
    if (P.d > 16 * QMAX_CHUNKS + 15 || P.max_depth + 2 > TB_STACK) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_iota_u32, dim3((P.n + 255) / 256), dim3(256), 0, s, P.perm, P.n);
    size_t lds = 2ull * P.d * sizeof(float);
    switch (P.d >> 4) {
        case 8: hipLaunchKernelGGL(k_tree_build<8>, dim3(1), dim3(TB_THREADS), lds, s, P); break;
        case 6: hipLaunchKernelGGL(k_tree_build<6>, dim3(1), dim3(TB_THREADS), lds, s, P); break;
        default: hipLaunchKernelGGL(k_tree_build<0>, dim3(1), dim3(TB_THREADS), lds, s, P); break;
    }
    return hipGetLastError();
}

hipError_t launch_tree_entries(const uint32_t* left, const uint32_t* leaf_start, const uint32_t* leaf_size,
                               const uint32_t* perm, uint32_t nodes, uint32_t* entry, hipStream_t s) {
    hipLaunchKernelGGL(k_tree_entries, dim3((nodes + 255) / 256), dim3(256), 0, s, left, leaf_start, leaf_size,
                       perm, nodes, entry);
    return hipGetLastError();
}
