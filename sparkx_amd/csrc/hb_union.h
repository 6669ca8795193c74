// hb_union.h -- the per-node HyperBall union step shared by the single-graph
// kernels (hb_kernels.hip) and the batched kernels (hb_batch.hip).
//
// For node u (hyper_ball.rs:803-831): a = c_{t-1}(u) | c_{t-1}(v) for every
// neighbour v; est(a) against the cached est(c_{t-1}(u)); f64 accumulation when
// the estimate increased; c_t(u) = a into the output buffer; changed-bit when any
// register changed.  A team of T lanes runs one node: G = 2^P/16 lanes hold one
// sketch copy (16 registers each), S = T/G copies split the neighbour list.
#pragma once

#include "hb_device.h"

namespace hb {

// one graph's state as seen by one iteration
struct HbView {
    const uint32_t *off, *nbr;
    const uint8_t *in;
    uint8_t *out;
    float *est;
    double *dist, *inv;
    const uint32_t *bitsR;  // registers changed at t-1
    uint32_t *bitsW;        // registers changed at t
    double t, inv_t;
    bool use_frontier;
};

struct HbCounts {
    unsigned long long inc = 0, arcs = 0, nodes = 0, chg = 0;
};

__device__ __forceinline__ bool bit_test(const uint32_t *bits, uint32_t v)
{
    return (bits[v >> 5] >> (v & 31u)) & 1u;
}

// f64 accumulation without contraction (hyper_ball.rs:819-828)
__device__ __forceinline__ bool accumulate(const HbView &v, uint32_t u, float prev, float curr)
{
    const float d32 = __fsub_rn(curr, prev);
    const double diff = (double)d32;
    if (diff > 0.0) {
        v.dist[u] = __dadd_rn(v.dist[u], __dmul_rn(v.t, diff));
        v.inv[u] = __dadd_rn(v.inv[u], __dmul_rn(v.inv_t, diff));
        return true;
    }
    return false;
}

template <int P, int T>
__device__ __forceinline__ void union_node(const HbView &v, const HbTables &tb, uint32_t u, uint32_t tl,
                                           HbCounts &c)
{
    constexpr uint32_t M = 1u << P, G = M / 16, S = T / G;
    static_assert(T >= (int)G && T <= HB_WAVE, "team width");
    (void)M;
    const uint32_t cl = tl % G, slot = tl / G;
    const uint32_t e0 = v.off[u], deg = v.off[u + 1] - e0;
    const uint32_t *__restrict__ nb = v.nbr + e0;
    if (v.use_frontier) {
        uint32_t any = 0;
        for (uint32_t j = tl; j < deg; j += T) any |= bit_test(v.bitsR, nb[j]);
        if (hb_copy_or<T>(any) == 0) {
            // no neighbour changed: c_t(u) = c_{t-1}(u); keep both buffers complete
            if (slot == 0 && bit_test(v.bitsR, u))
                *(uint4 *)(v.out + ((uint64_t)u << P) + cl * 16) =
                    *(const uint4 *)(v.in + ((uint64_t)u << P) + cl * 16);
            return;
        }
    }
    const uint8_t *__restrict__ in = v.in;
    const uint4 own = *(const uint4 *)(in + ((uint64_t)u << P) + cl * 16);
    HbAcc pa = hb_acc(own);
    uint32_t j = slot;
    for (; j + 3 * S < deg; j += 4 * S) {
        const uint32_t v0 = nb[j], v1 = nb[j + S], v2 = nb[j + 2 * S], v3 = nb[j + 3 * S];
        const uint4 x0 = *(const uint4 *)(in + ((uint64_t)v0 << P) + cl * 16);
        const uint4 x1 = *(const uint4 *)(in + ((uint64_t)v1 << P) + cl * 16);
        const uint4 x2 = *(const uint4 *)(in + ((uint64_t)v2 << P) + cl * 16);
        const uint4 x3 = *(const uint4 *)(in + ((uint64_t)v3 << P) + cl * 16);
        hb_acc_max(pa, x0);
        hb_acc_max(pa, x1);
        hb_acc_max(pa, x2);
        hb_acc_max(pa, x3);
    }
    for (; j < deg; j += S) hb_acc_max(pa, *(const uint4 *)(in + ((uint64_t)nb[j] << P) + cl * 16));
    uint4 acc = hb_acc_get(pa);
#pragma unroll
    for (uint32_t s = G; s < T; s <<= 1) acc = hb_bmax4(acc, hb_shfl_xor4<T>(acc, s));

    const uint32_t diffw = (acc.x != own.x) | (acc.y != own.y) | (acc.z != own.z) | (acc.w != own.w);
    const bool changed = hb_copy_or<G>(diffw) != 0;
    const float curr = hb_estimate_copy<G>(tb, acc);
    if (slot == 0) *(uint4 *)(v.out + ((uint64_t)u << P) + cl * 16) = acc;
    if (tl == 0) {
        const float prev = v.est[u];
        if (accumulate(v, u, prev, curr)) c.inc += 1;
        v.est[u] = curr;
        if (changed) {
            atomicOr(v.bitsW + (u >> 5), 1u << (u & 31u));
            c.chg += 1;
        }
        c.arcs += deg;
        c.nodes += 1;
    }
}

}  // namespace hb
