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
    const uint32_t *orig;      // hashed id of each arc's head (iteration 1)
    const uint32_t *node_id;   // hashed id of each node (null: its index)
    const uint8_t *in;
    bool first;             // iteration 1 with P <= 10: sketches hold only their own insertion
    uint32_t *scr;          // LDS, 16 u32 per lane of the block (iteration 1 only)
    uint8_t *out;
    float *est;
    double *dist, *inv;
    const uint32_t *bitsR;  // registers changed at t-1
    const uint32_t *gather_bits;  // bitsR when only changed neighbours need gathering, else null
    uint32_t *bitsW;        // registers changed at t
    double t, inv_t;
    bool use_frontier;
};

// One lane's counts over its share of a launch.  32 bits suffice (a lane's sums are at
// most the graph's node or arc count, both below 2^32 in the u32 CSR) and keep five
// registers free in the union kernels; the block and batch sums widen them to 64 bits.
struct HbCounts {
    uint32_t inc = 0, arcs = 0, nodes = 0, chg = 0;
    uint32_t gathered = 0;  // neighbour register blocks read
};

__device__ __forceinline__ bool bit_test(const uint32_t *bits, uint32_t v)
{
    return (bits[v >> 5] >> (v & 31u)) & 1u;
}

// f64 accumulation without contraction (hyper_ball.rs:819-828); dist0 / inv0 are the
// node's accumulators as loaded by the caller
__device__ __forceinline__ bool accumulate(const HbView &v, uint32_t u, float prev, float curr, double dist0,
                                           double inv0)
{
    const float d32 = __fsub_rn(curr, prev);
    const double diff = (double)d32;
    if (diff > 0.0) {
        v.dist[u] = __dadd_rn(dist0, __dmul_rn(v.t, diff));
        v.inv[u] = __dadd_rn(inv0, __dmul_rn(v.inv_t, diff));
        return true;
    }
    return false;
}
__device__ __forceinline__ bool accumulate(const HbView &v, uint32_t u, float prev, float curr)
{
    const float d32 = __fsub_rn(curr, prev);
    if (!((double)d32 > 0.0)) return false;  // the accumulators are read only when needed
    return accumulate(v, u, prev, curr, v.dist[u], v.inv[u]);
}

// (index << 6) | rank of the insertion of node id v at precision P (P <= 10)
template <int P>
__device__ __forceinline__ uint32_t hb_insertion(uint32_t v)
{
    uint32_t idx, rank;
    hb_index_rank(v, P, &idx, &rank);
    return (idx << 6) | rank;
}

// This lane's 16 registers of c_0(v) for copy lane cl, from w = (index << 6) | rank.
__device__ __forceinline__ uint4 hb_own_from_word(uint32_t w, uint32_t cl)
{
    const uint32_t idx = w >> 6, rank = w & 63u;
    uint4 r = make_uint4(0, 0, 0, 0);
    if ((idx >> 4) == cl) {
        const uint32_t v = rank << (8u * (idx & 3u)), q = (idx >> 2) & 3u;
        r.x = q == 0 ? v : 0u;
        r.y = q == 1 ? v : 0u;
        r.z = q == 2 ? v : 0u;
        r.w = q == 3 ? v : 0u;
    }
    return r;
}

// register idx of the copy := max(register, rank) when it falls in lane cl's 16
__device__ __forceinline__ void hb_put_max(uint4 &a, uint32_t w, uint32_t cl)
{
    const uint32_t idx = w >> 6;
    if ((idx >> 4) != cl) return;
    const uint32_t sh = 8u * (idx & 3u), q = (idx >> 2) & 3u;
    const uint32_t m = 0xffu << sh, r = (w & 63u) << sh;
    uint32_t x = q == 0 ? a.x : q == 1 ? a.y : q == 2 ? a.z : a.w;
    if ((x & m) < r) x = (x & ~m) | r;
    if (q == 0) a.x = x;
    else if (q == 1) a.y = x;
    else if (q == 2) a.z = x;
    else a.w = x;
}

// Union of the sketches of nb[slot], nb[slot + S], ... (j < len) into `start`,
// for the 16 registers of copy lane cl (the copy is G = 2^P/16 consecutive lanes).
// With `changed` (the nodes whose registers changed at t-1) only those neighbours are
// gathered: an unchanged neighbour v has c_{t-1}(v) = c_{t-2}(v), which c_{t-1}(u)
// already contains (v is u's neighbour), so leaving it out changes nothing.
// In iteration 1 (`first`) every neighbour's sketch is still its own insertion, so
// instead of gathering its 2^P-byte register block the copy's lanes hash the
// neighbour ids, one each, and raise register idx to the rank with an LDS atomic max
// on a u32 per register (`scr`: 16 u32 per lane, the copy's registers contiguous
// from the copy's first lane); no random global memory access at all.
template <int P, uint32_t S, bool SEL>
__device__ __forceinline__ uint4 hb_gather_union(const uint8_t *__restrict__ in, bool first, uint32_t *scr,
                                                 const uint32_t *__restrict__ changed,
                                                 const uint32_t *__restrict__ nb,
                                                 const uint32_t *__restrict__ nbh, uint32_t len, uint32_t slot,
                                                 uint32_t cl, uint4 start, uint32_t &gathered)
{
    constexpr uint32_t G = (1u << P) / 16;
    uint32_t j = slot;
    if (first) {
        uint32_t *mine = scr + threadIdx.x * 16u;              // this lane's 16 registers
        uint32_t *copy = scr + (threadIdx.x - cl) * 16u;       // register idx of the copy at copy[idx]
        const uint32_t sw[4] = {start.x, start.y, start.z, start.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) mine[k] = (sw[k >> 2] >> (8 * (k & 3))) & 0xffu;
        __builtin_amdgcn_wave_barrier();
        for (uint32_t q = slot + cl * S; q < len; q += G * S) {
            const uint32_t w = hb_insertion<P>(nbh[q]);
            atomicMax(copy + (w >> 6), w & 63u);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        uint32_t r[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
            r[k] = mine[4 * k] | (mine[4 * k + 1] << 8) | (mine[4 * k + 2] << 16) | (mine[4 * k + 3] << 24);
        return make_uint4(r[0], r[1], r[2], r[3]);
    }
    HbAcc pa = hb_acc(start);
    if constexpr (SEL) {
        const uint4 z = make_uint4(0, 0, 0, 0);  // registers are >= 0: max with 0 is a no-op
        for (; j + 3 * S < len; j += 4 * S) {
            const uint32_t v0 = nb[j], v1 = nb[j + S], v2 = nb[j + 2 * S], v3 = nb[j + 3 * S];
            const bool c0 = bit_test(changed, v0), c1 = bit_test(changed, v1);
            const bool c2 = bit_test(changed, v2), c3 = bit_test(changed, v3);
            gathered += (uint32_t)c0 + (uint32_t)c1 + (uint32_t)c2 + (uint32_t)c3;
            const uint4 x0 = c0 ? *(const uint4 *)(in + ((uint64_t)v0 << P) + cl * 16) : z;
            const uint4 x1 = c1 ? *(const uint4 *)(in + ((uint64_t)v1 << P) + cl * 16) : z;
            const uint4 x2 = c2 ? *(const uint4 *)(in + ((uint64_t)v2 << P) + cl * 16) : z;
            const uint4 x3 = c3 ? *(const uint4 *)(in + ((uint64_t)v3 << P) + cl * 16) : z;
            hb_acc_max(pa, x0);
            hb_acc_max(pa, x1);
            hb_acc_max(pa, x2);
            hb_acc_max(pa, x3);
        }
        for (; j < len; j += S) {
            const uint32_t v = nb[j];
            if (bit_test(changed, v)) {
                hb_acc_max(pa, *(const uint4 *)(in + ((uint64_t)v << P) + cl * 16));
                ++gathered;
            }
        }
        return hb_acc_get(pa);
    } else {
        (void)changed;
    }
    gathered += len > slot ? (len - slot + S - 1) / S : 0;
    for (; j + 3 * S < len; j += 4 * S) {
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
    for (; j < len; j += S) hb_acc_max(pa, *(const uint4 *)(in + ((uint64_t)nb[j] << P) + cl * 16));
    return hb_acc_get(pa);
}

// SEL: gather only the neighbours in v.gather_bits (a separate instantiation, so the
// dense path's code is unchanged by the selective one)
template <int P, int T, bool SEL = false>
__device__ __forceinline__ void union_node(const HbView &v, const HbTables &tb, uint32_t u, uint32_t tl,
                                           HbCounts &c)
{
    constexpr uint32_t M = 1u << P, G = M / 16, S = T / G;
    static_assert(T >= (int)G && T <= HB_WAVE, "team width");
    (void)M;
    const uint32_t cl = tl % G, slot = tl / G;
    const uint32_t e0 = v.off[u], deg = v.off[u + 1] - e0;
    const uint32_t *__restrict__ nb = v.nbr + e0;
    const uint32_t *__restrict__ nbh = v.orig + e0;
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
    const uint4 own = v.first ? hb_own_from_word(hb_insertion<P>(v.node_id ? v.node_id[u] : u), cl)
                              : *(const uint4 *)(v.in + ((uint64_t)u << P) + cl * 16);
    uint32_t gathered = 0;  // by this lane's copy
    uint4 acc = hb_gather_union<P, S, SEL>(v.in, v.first, v.scr, v.gather_bits, nb, nbh, deg, slot, cl, own,
                                                   gathered);
    if (cl == 0) c.gathered += gathered;
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

// union_node for a dense iteration (every neighbour gathered, no frontier check) with one
// sketch copy per team (T = G), where the wavefront's 64 / G teams hold the consecutive
// nodes lo .. hi - 1 (a bin that is an id range).  Each team puts its own row into the
// wavefront's LDS window `win` (64 uint4, copy lane cl of node w at (w - lo) G + cl), and a
// neighbour inside [lo, hi) is read from there instead of memory: in a numbering that
// keeps chains together (the chain layout) most neighbours are in the window, and the
// union is bound by the rate of memory requests, not by bytes.  Same registers,
// estimates and accumulators as union_node.  (At p = 5 and 6 the single-graph engine runs
// such iterations with one node per lane instead: k_union_rows.)
template <int P, int T>
__device__ __forceinline__ void union_node_window(const HbView &v, const HbTables &tb, uint32_t u, uint32_t tl,
                                                  HbCounts &c, uint32_t lo, uint32_t hi, uint4 *win)
{
    constexpr uint32_t G = (1u << P) / 16;
    static_assert(T == (int)G, "one copy per team");
    const uint32_t cl = tl;
    const uint32_t e0 = v.off[u], deg = v.off[u + 1] - e0;
    const uint32_t *__restrict__ nb = v.nbr + e0;
    const uint8_t *__restrict__ in = v.in;
    // the node's estimate is loaded beside its own row, its accumulators right after the
    // gathers: each then arrives while the union and the estimate are computed instead of
    // costing a memory round trip of its own after them (≈ 1% on the k-mer graph)
    const float prev = tl == 0 ? v.est[u] : 0.0f;
    const uint4 own = *(const uint4 *)(in + ((uint64_t)u << P) + cl * 16);
    win[(u - lo) * G + cl] = own;
    // LDS operations of one wavefront complete in order; the barrier and fence keep the
    // compiler from moving the window reads above the write
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const uint32_t span = hi - lo;
    HbAcc pa = hb_acc(own);
    uint32_t j = 0;
    for (; j + 1 < deg; j += 2) {
        const uint32_t w0 = nb[j], w1 = nb[j + 1];
        const uint4 x0 = w0 - lo < span ? win[(w0 - lo) * G + cl] : *(const uint4 *)(in + ((uint64_t)w0 << P) + cl * 16);
        const uint4 x1 = w1 - lo < span ? win[(w1 - lo) * G + cl] : *(const uint4 *)(in + ((uint64_t)w1 << P) + cl * 16);
        hb_acc_max(pa, x0);
        hb_acc_max(pa, x1);
    }
    if (j < deg) {
        const uint32_t w0 = nb[j];
        hb_acc_max(pa, w0 - lo < span ? win[(w0 - lo) * G + cl] : *(const uint4 *)(in + ((uint64_t)w0 << P) + cl * 16));
    }
    const uint4 acc = hb_acc_get(pa);
    double dist0 = 0.0, inv0 = 0.0;
    if (tl == 0) {
        dist0 = v.dist[u];
        inv0 = v.inv[u];
    }
    if (cl == 0) c.gathered += deg;
    // every team has read the window before the next trip overwrites it
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");

    const uint32_t diffw = (acc.x != own.x) | (acc.y != own.y) | (acc.z != own.z) | (acc.w != own.w);
    const bool changed = hb_copy_or<G>(diffw) != 0;
    const float curr = hb_estimate_copy<G>(tb, acc);
    *(uint4 *)(v.out + ((uint64_t)u << P) + cl * 16) = acc;
    if (tl == 0) {
        if (accumulate(v, u, prev, curr, dist0, inv0)) c.inc += 1;
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
