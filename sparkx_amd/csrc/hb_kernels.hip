// hb_kernels.hip -- HyperBall iteration kernels for gfx950 (no MFMA: the union
// is a bytewise max, HBM-bound).
//
// One iteration t (hyper_ball.rs:754-874, restated per node u):
//     a = c_{t-1}(u); for v in N(u): a |= c_{t-1}(v)       (register-wise max)
//     diff = f64(est(a) - est(c_{t-1}(u)))  (f32 subtraction)
//     if diff > 0 { dist[u] += t*diff; inv[u] += (1/t)*diff; increases++ }
//     c_t(u) = a
// Launch sequence per iteration, all reading the device-resident HbIterState so
// that one captured hipGraph serves every iteration:
//     k_begin (clear next changed-bitmap)  ->  k_union<P,T> per degree bin
//     (k_union_rows<P> for the dense iterations of a chain-layout id range at p = 5, 6)
//     ->  k_chunk<P> (hub neighbour chunks) -> k_merge<P> (hubs) -> [k_nf] -> k_finalize
// Frontier: if no neighbour of u changed at t-1 then c_t(u) = c_{t-1}(u) (union
// of unchanged sets), so u is skipped; when u itself changed at t-1 its row is
// copied into the output buffer so both buffers stay complete.  Results are
// identical to recomputing every node.
#include <cmath>

#include "hb_device.h"
#include "hb_kernels.h"
#include "hb_union.h"

namespace hb {

static constexpr int kBlock = 256;


// block-wide sums of per-lane counters, one atomic per counter per block
__device__ __forceinline__ void block_add5(unsigned long long a, unsigned long long b, unsigned long long c,
                                           unsigned long long e, unsigned long long f, HbIterRec *rec, int slot)
{
    __shared__ unsigned long long red[5][kBlock / HB_WAVE];
#pragma unroll
    for (int s = 1; s < HB_WAVE; s <<= 1) {
        a += __shfl_xor(a, s);
        b += __shfl_xor(b, s);
        c += __shfl_xor(c, s);
        e += __shfl_xor(e, s);
        f += __shfl_xor(f, s);
    }
    const int w = threadIdx.x / HB_WAVE;
    if ((threadIdx.x & (HB_WAVE - 1)) == 0) {
        red[0][w] = a;
        red[1][w] = b;
        red[2][w] = c;
        red[3][w] = e;
        red[4][w] = f;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
        for (int i = 0; i < (int)(blockDim.x / HB_WAVE); ++i) {
            s0 += red[0][i];
            s1 += red[1][i];
            s2 += red[2][i];
            s3 += red[3][i];
            s4 += red[4][i];
        }
        if (s0) atomicAdd(&rec->inc[slot], s0);
        if (s1) atomicAdd(&rec->arcs[slot], s1);
        if (s2) atomicAdd(&rec->nodes[slot], s2);
        if (s3) atomicAdd(&rec->chg[slot], s3);
        if (s4) atomicAdd(&rec->gathered[slot], s4);
    }
}

// f64 accumulation without contraction (hyper_ball.rs:823-825)
__device__ __forceinline__ bool accumulate(const HbDev &d, uint32_t u, float prev, float curr,
                                           double t, double inv_t)
{
    const float d32 = __fsub_rn(curr, prev);
    const double diff = (double)d32;
    if (diff > 0.0) {
        d.dist[u] = __dadd_rn(d.dist[u], __dmul_rn(t, diff));
        d.inv[u] = __dadd_rn(d.inv[u], __dmul_rn(inv_t, diff));
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// init: c_0(u) = {insert(u)} (hyper_ball.rs:716-721); est, accumulators
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(kBlock) void k_init(HbDev d)
{
    // One thread per node computes the node's insertion (one SipHash-1-3 per node); the
    // wavefront then writes the rows of its 64 nodes, G lanes per row and 64 / G rows per
    // store, taking each row's insertion by shuffle, so the stores stay coalesced.
    constexpr uint32_t M = 1u << P, G = M / 16, RPI = HB_WAVE / G;
    const uint32_t lane = threadIdx.x % HB_WAVE, cl = lane % G;
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t first = u - lane;  // the wavefront's first node
    uint32_t word = 0;                // rank | idx << 8 | (no neighbours) << 31
    if (u < d.n) {
        uint32_t idx, rank;
        hb_index_rank(d.node_id ? d.node_id[u] : (uint32_t)u, P, &idx, &rank);
        word = rank | (idx << 8) | (d.off[u + 1] == d.off[u] ? 1u << 31 : 0u);
    }
#pragma unroll
    for (uint32_t k = 0; k < G; ++k) {
        const uint32_t r = k * RPI + lane / G;  // this store's row among the wavefront's nodes
        const uint32_t w = (uint32_t)__shfl((int)word, (int)r);
        const uint64_t node = first + r;
        if (node < d.n) {
            const uint32_t idx = (w >> 8) & 0xffffu, rank = w & 0xffu;
            uint32_t q[4] = {0, 0, 0, 0};
            if (idx / 16 == cl) q[(idx % 16) / 4] = rank << (8 * (idx % 4));
            const uint4 row = make_uint4(q[0], q[1], q[2], q[3]);
            *(uint4 *)(d.regs0 + (node << P) + cl * 16) = row;
            // a node without neighbours keeps c_0 forever and is never scheduled: both buffers
            if (w >> 31) *(uint4 *)(d.regs1 + (node << P) + cl * 16) = row;
        }
    }
    if (u < d.n) {
        // c_0 holds one register of rank >= 1 and m - 1 empty ones, so the estimator returns
        // linear counting of m - 1 zeros (launch_init checks that the HLL++ threshold admits it)
        d.est[u] = d.lc[M - 1];
        d.dist[u] = 0.0;
        d.inv[u] = 0.0;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t *p, uint64_t count, uint32_t value)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---------------------------------------------------------------------------
// push iterations (symmetric graph, few changes at t-1): u must be recomputed
// iff one of its neighbours changed, i.e. iff u is a neighbour of a changed node.
// k_begin sets those bits from the changed nodes' lists (a wavefront per bitmap
// word); k_push_compact lists them per degree bin with wavefront-aggregated appends
// and copies c_{t-1}(u) into the output buffer for every non-hub node that changed
// at t-1 but has no changed neighbour (hubs: k_merge copies).
// ---------------------------------------------------------------------------
// Start of iteration t: clear the bitmap iteration t writes (bits[t & 1]) and, where
// push iterations can occur, the active bitmap of iteration t + 1; clear this
// iteration's ring record; in a push iteration also clear the per-bin list counts and
// mark this iteration's active bitmap.  The marking reads bits[(t - 1) & 1] and writes
// hb_active(t), neither of which the clearing touches.
__global__ __launch_bounds__(kBlock) void k_begin(HbDev d, uint32_t words)
{
    const HbIterState *st = d.st;
    if (st->done) return;
    const uint32_t it = st->iter;
    uint32_t *bw = (it & 1u) ? d.bits1 : d.bits0;
    const bool push = st->push != 0;
    if (push && blockIdx.x == 0 && threadIdx.x < 3) d.dyn_count[threadIdx.x] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {  // this iteration's ring record, cleared
        HbIterRec *r = d.rec + it % HB_REC_RING;
        *r = HbIterRec{};
        if (push) r->push = 1ull;
    }
    uint32_t *next_active = d.active ? hb_active(d, it + 1) : nullptr;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) {
        bw[i] = 0u;
        if (next_active) next_active[i] = 0u;
    }
    if (!push) return;
    uint32_t *act = hb_active(d, it);
    const uint32_t *bitsR = (it & 1u) ? d.bits0 : d.bits1;
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1);
    const uint32_t waves = gridDim.x * (kBlock / HB_WAVE);
    for (uint32_t wd = blockIdx.x * (kBlock / HB_WAVE) + threadIdx.x / HB_WAVE; wd < words; wd += waves) {
        uint32_t m = bitsR[wd];  // one word per wavefront: the loop below is uniform
        while (m) {
            const uint32_t w = wd * 32u + (uint32_t)(__ffs(m) - 1);
            m &= m - 1u;
            const uint32_t e1 = d.off[w + 1];
            for (uint32_t e = d.off[w] + lane; e < e1; e += HB_WAVE) {
                const uint32_t u = d.nbr[e], bit = 1u << (u & 31u);
                if (!(act[u >> 5] & bit)) atomicOr(act + (u >> 5), bit);
            }
        }
    }
}

template <int P>
__global__ __launch_bounds__(kBlock) void k_push_compact(HbDev d, uint32_t words)
{
    const HbIterState *st = d.st;
    if (st->done || !st->push) return;
    const uint32_t *bitsR = (st->iter & 1u) ? d.bits0 : d.bits1;
    const uint8_t *in = st->parity ? d.regs1 : d.regs0;
    uint8_t *out = st->parity ? d.regs0 : d.regs1;
    const uint32_t *act = hb_active(d, st->iter);
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1);
    // the trip count is uniform across a wavefront, so every lane takes part in the ballots
    for (uint32_t base = blockIdx.x * blockDim.x; base < words; base += gridDim.x * blockDim.x) {
        const uint32_t wd = base + threadIdx.x;
        uint32_t a = 0, c = 0;
        if (wd < words) {
            a = act[wd];
            c = bitsR[wd] & ~a;
        }
        while (c) {
            const uint32_t u = wd * 32u + (uint32_t)(__ffs(c) - 1);
            c &= c - 1u;
            if (d.off[u + 1] - d.off[u] > d.wave_max) continue;
            const uint4 *src = (const uint4 *)(in + ((uint64_t)u << P));
            uint4 *dst = (uint4 *)(out + ((uint64_t)u << P));
#pragma unroll
            for (int k = 0; k < (1 << P) / 16; ++k) dst[k] = src[k];
        }
        while (__ballot(a != 0u)) {
            uint32_t u = 0, bin = 3;
            if (a) {
                u = wd * 32u + (uint32_t)(__ffs(a) - 1);
                a &= a - 1u;
                const uint32_t deg = d.off[u + 1] - d.off[u];
                // hubs (bin 3) are not listed: k_chunk / k_merge check their own chunks;
                // deg 0 cannot occur on a symmetric graph and is never listed either
                bin = deg == 0 ? 3u : deg <= d.small_max ? 0u : deg <= d.mid_max ? 1u : deg <= d.wave_max ? 2u : 3u;
            }
            for (uint32_t b = 0; b < 3; ++b) {
                const unsigned long long m = __ballot(bin == b);
                if (!m) continue;
                const int leader = __ffsll(m) - 1;
                uint32_t pos = 0;
                if ((int)lane == leader) pos = atomicAdd(d.dyn_count + b, (uint32_t)__popcll(m));
                pos = __shfl(pos, leader);
                if (bin == b)
                    d.dyn_list[d.dyn_base[b] + pos + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = u;
            }
        }
    }
}

// Order-free fingerprints of the arc multiset and of its reverse, with h a
// splitmix64 mix of the pair: acc[0] = sum h(u, v), acc[1] = sum h(v, u),
// acc[2] = sum h(u, v)^2, acc[3] = sum h(v, u)^2 (wrapping).  A symmetric multiset
// gives acc[0] == acc[1] and acc[2] == acc[3]; any other CSR matches both sums only
// by a coincidence of two 64-bit equations.
__device__ __forceinline__ uint64_t hb_mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// A wavefront takes 64 consecutive nodes at a time: a node of degree <= 64 is walked
// by its own lane, a larger one by the whole wavefront (ballot over the 64, then a
// uniform loop), so hubs do not serialise on one lane.  Each block writes its four
// sums to part[4 * block]; sum_columns adds the blocks in a fixed order.
__device__ __forceinline__ void hb_fp_add(unsigned long long *a, uint64_t u, uint64_t v)
{
    const uint64_t hf = hb_mix64((u << 32) | v), hb = hb_mix64((v << 32) | u);
    a[0] += hf;
    a[1] += hb;
    a[2] += hf * hf;
    a[3] += hb * hb;
}

// every thread takes kFpArcs consecutive arcs: a binary search over the offsets finds
// the first one's tail, then the walk advances the tail across node boundaries (empty
// nodes between two arcs are crossed by exactly one thread), so hubs cost no more than
// any other arcs
constexpr uint32_t kFpArcs = 32;

__global__ __launch_bounds__(kBlock) void k_arc_fingerprint(const uint32_t *off, const uint32_t *nbr, uint32_t n,
                                                            uint64_t arcs, unsigned long long *part)
{
    __shared__ unsigned long long red[kBlock / HB_WAVE][4];
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1), wv = threadIdx.x / HB_WAVE;
    unsigned long long a[4] = {0, 0, 0, 0};
    const uint64_t threads = (uint64_t)gridDim.x * kBlock;
    for (uint64_t t = (uint64_t)blockIdx.x * kBlock + threadIdx.x; t * kFpArcs < arcs; t += threads) {
        const uint64_t e_lo = t * kFpArcs, e_hi = min<uint64_t>(arcs, e_lo + kFpArcs);
        uint32_t lo = 0, hi = n - 1;  // largest u with off[u] <= e_lo (then off[u + 1] > e_lo)
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo + 1) / 2;
            if (off[mid] <= e_lo) lo = mid;
            else hi = mid - 1;
        }
        uint32_t u = lo;
        for (uint64_t e = e_lo; e < e_hi; ++e) {
            while (off[u + 1] <= e) ++u;
            hb_fp_add(a, u, nbr[e]);
        }
    }
    for (int k = 0; k < 4; ++k)
        for (int o = 1; o < HB_WAVE; o <<= 1) a[k] += __shfl_xor(a[k], o);
    if (lane == 0)
        for (int k = 0; k < 4; ++k) red[wv][k] = a[k];
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned long long sum = 0;
        for (int w = 0; w < kBlock / HB_WAVE; ++w) sum += red[w][threadIdx.x];
        part[4 * blockIdx.x + threadIdx.x] = sum;
    }
}

// ---------------------------------------------------------------------------
// dense iterations of an id-range bin in the chain layout at p = 5 or 6: one lane per
// node, its whole 2^p-byte row in the lane's registers (W = 2^p / 16 words)
// ---------------------------------------------------------------------------
// With a team of G = 2^p / 16 lanes per node (union_node_window) a wavefront holds 16
// nodes at p = 6, and the lanes of a team repeat each other's per-node work: offsets,
// addresses, the estimate.  On a graph of long low-degree chains that union is bound by
// each node's chain of dependent memory accesses, not by bytes, so what counts is how
// many nodes are in flight.  Here a wavefront holds 64 consecutive nodes lo..lo+63:
//   1. their rows c_{t-1} enter the wavefront's LDS window by coalesced 16-byte loads;
//   2. each lane takes its node's row from the window and merges its neighbours' rows,
//      from the window when the neighbour is one of the 64, else from memory;
//   3. the new rows go back through the window and out by coalesced stores;
//   4. estimate and accumulators per lane; the changed bits of the 64 nodes are ORed
//      into the bitmap with one atomic per 32-bit word.
// Same registers, estimates and accumulators as union_node.

// an iteration whose id-range bins k_union_rows runs instead of k_union: dense (no
// frontier gathers, no push list).  That includes the first: k_init wrote every node's
// c_0 row into the buffer iteration 1 reads, and gathering those rows, most of them from
// the LDS window, is cheaper here than hashing every neighbour id as k_union does.
__device__ __forceinline__ bool rows_iteration(const HbDev &d, const HbIterState &st)
{
    return !st.push && (st.iter == 1 || !(d.frontier && !st.skip_scan));
}

// The window of a wavefront is W planes of 64 uint4: plane q holds chunk q (bytes
// 16q .. 16q + 15) of the 64 rows, row w at position w ^ row_perm(q).  The permutation
// keeps a wavefront's structured accesses free of LDS bank conflicts: a lane reading
// chunk q of its own row, and the coalesced copies (chunk q of rows 64/W apart per
// lane group), land in distinct 16-byte bank groups for 8 and for 16 lanes per cycle.
template <int W>
__device__ __forceinline__ uint32_t row_perm(uint32_t q)
{
    return W == 4 ? (0xCA60u >> (4 * q)) & 15u : (0xC0u >> (4 * q)) & 15u;  // {0,6,10,12} / {0,12}
}

template <int W>
__device__ __forceinline__ uint32_t win_index(uint32_t w, uint32_t q)
{
    return q * HB_WAVE + (w ^ row_perm<W>(q));
}

// row c_{t-1}(x): from the window when x is one of its nodes lo .. lo + span - 1
template <int P>
__device__ __forceinline__ void row_fetch(uint4 *r, uint32_t x, uint32_t lo, uint32_t span, const uint4 *win,
                                          const uint8_t *__restrict__ in)
{
    constexpr int W = (1 << P) / 16;
    if (x - lo < span) {
#pragma unroll
        for (int q = 0; q < W; ++q) r[q] = win[win_index<W>(x - lo, q)];
    } else {
        const uint4 *g = (const uint4 *)(in + ((uint64_t)x << P));
#pragma unroll
        for (int q = 0; q < W; ++q) r[q] = g[q];
    }
}

// Seven wavefronts per SIMD (71 VGPRs at p = 6, no spill): 448 nodes in flight per SIMD
// against 128 for union_node_window at eight wavefronts of 16.  On the k-mer graph (p = 6) a
// launch takes 0.362 ms against 0.583 for union_node_window; at eight (64 VGPRs, spills
// outside the node loop only) 0.365 ms.  The bin's base must be a multiple of 64.
template <int P>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(7))) void k_union_rows(
    HbDev d, uint32_t base, uint32_t count, int slot_id)
{
    constexpr int W = (1 << P) / 16;
    static_assert(W == 2 || W == 4, "p = 5 or 6");
    const HbIterState *st = d.st;
    if (st->done || !rows_iteration(d, *st)) return;
    HB_TABLES_LDS(P, d, tb)
    __shared__ uint4 s_win[kBlock * W];  // 64 rows per wavefront
    const uint32_t it = st->iter;
    const uint32_t lane = threadIdx.x % HB_WAVE;
    uint4 *win = s_win + (threadIdx.x / HB_WAVE) * (HB_WAVE * W);
    HbView v{};
    v.in = st->parity ? d.regs1 : d.regs0;
    v.out = st->parity ? d.regs0 : d.regs1;
    v.dist = d.dist;
    v.inv = d.inv;
    v.t = st->t;
    v.inv_t = st->inv_t;
    const uint8_t *__restrict__ in = v.in;
    const uint32_t *__restrict__ nb = d.nbr;
    uint32_t *bitsW = (it & 1u) ? d.bits1 : d.bits0;
    // chunk r * 64 + lane of the wavefront's rows (a coalesced copy) is chunk lane % W of
    // row r * 64 / W + lane / W, at window index cidx + r * 64 / W (the permutation only
    // touches the low 4 bits of a row, below 64 / W)
    constexpr uint32_t RPC = HB_WAVE / W;  // rows per coalesced copy instruction
    const uint32_t lw = lane / W;
    const uint32_t cidx = win_index<W>(lw, lane % W);
    unsigned long long n_inc = 0, n_arcs = 0, n_nodes = 0, n_chg = 0;  // the wavefront's counts
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / HB_WAVE);
    const uint32_t stride = __builtin_amdgcn_readfirstlane(gridDim.x) * kBlock;
    for (uint32_t w0 = (blockIdx.x * (kBlock / HB_WAVE) + wave) * HB_WAVE; w0 < count; w0 += stride) {
        const uint32_t lo = base + w0, span = min((uint32_t)HB_WAVE, count - w0);
        const bool valid = lane < span;
        const uint32_t u = lo + lane;
        uint32_t e0 = 0, e1 = 0;
        float prev = 0.0f;
        if (valid) {
            e0 = d.off[u];
            e1 = d.off[u + 1];
            prev = d.est[u];
        }
        const uint4 *src = (const uint4 *)(in + ((uint64_t)lo << P));
#pragma unroll
        for (int r = 0; r < W; ++r)
            if ((int)lw < (int)span - r * (int)RPC) win[cidx + r * RPC] = src[r * HB_WAVE + lane];
        // LDS operations of one wavefront complete in order; the barrier and fence keep
        // the compiler from moving window accesses across each other
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        uint4 acc[W];
#pragma unroll
        for (int q = 0; q < W; ++q) acc[q] = valid ? win[win_index<W>(lane, q)] : make_uint4(0, 0, 0, 0);
        uint32_t e = e0;
        for (; e < e1; ++e) {
            uint4 r0[W];
            row_fetch<P>(r0, nb[e], lo, span, win, in);
#pragma unroll
            for (int q = 0; q < W; ++q) acc[q] = hb_bmax4(acc[q], r0[q]);
        }
        __builtin_amdgcn_wave_barrier();  // every lane has read the window
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        uint32_t diffw = 0;
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const uint4 o = win[win_index<W>(lane, q)];
            diffw |= (acc[q].x != o.x) | (acc[q].y != o.y) | (acc[q].z != o.z) | (acc[q].w != o.w);
            win[win_index<W>(lane, q)] = acc[q];
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        uint4 *dst = (uint4 *)(v.out + ((uint64_t)lo << P));
#pragma unroll
        for (int r = 0; r < W; ++r)
            if ((int)lw < (int)span - r * (int)RPC) dst[r * HB_WAVE + lane] = win[cidx + r * RPC];
        // counts: nodes lo .. lo + span - 1 hold arcs off[lo] .. off[lo + span] - 1
        n_nodes += span;
        n_arcs += (uint32_t)__builtin_amdgcn_readlane((int)e1, (int)span - 1) -
                  (uint32_t)__builtin_amdgcn_readfirstlane((int)e0);
        const float curr = hb_estimate_row<W>(tb, acc);
        const bool changed = valid && diffw != 0;
        bool inc = false;
        if (valid) {
            inc = accumulate(v, u, prev, curr);  // reads the accumulators only if they change
            d.est[u] = curr;
        }
        n_inc += __popcll(__ballot(inc));
        // the changed bits of nodes lo .. lo + 63 (lo a multiple of 64): two bitmap words,
        // ORed atomically since the last may also hold bits of nodes past the bin
        const unsigned long long mask = __ballot(changed);
        n_chg += __popcll(mask);
        const uint32_t bits_lo = (uint32_t)mask, bits_hi = (uint32_t)(mask >> 32);
        if (lane == 0 && bits_lo) atomicOr(bitsW + (lo >> 5), bits_lo);
        if (lane == 1 && bits_hi) atomicOr(bitsW + (lo >> 5) + 1, bits_hi);
        // the next trip's window writes follow this trip's reads
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    const bool first_lane = lane == 0;
    block_add5(first_lane ? n_inc : 0, first_lane ? n_arcs : 0, first_lane ? n_nodes : 0, first_lane ? n_chg : 0,
               first_lane ? n_arcs : 0, d.rec + it % HB_REC_RING, slot_id);
}

// ---------------------------------------------------------------------------
// union over the neighbours of listed nodes; a team of T lanes per node
// (T = G: one sketch copy; T > G: S = T/G copies split the neighbours)
// ---------------------------------------------------------------------------
// Wavefronts per SIMD: the gathers are latency-bound, so more wavefronts in flight
// help until the register bound makes the compiler spill.  8 (at most 64 VGPRs) took the
// k-mer graph (p = 6, four-lane teams in the degree layout) from 59.5 to 49.5 ms per run.
// At p >= 7 the one-copy bins run at 8 too since their per-lane counters are 32-bit
// (HbCounts): 64 VGPRs without a spill, R-MAT-24's degree <= 256 bin 2.145 -> 2.130 ms per
// launch against 7 (with 64-bit counters 8 spilled: 41.8 ms per run against 41.0 at 7).
// The teams of 32 and 64 lanes spill 10-12 VGPRs at 72 (6 with the 32-bit counters); at 6
// (80 VGPRs) they do not, and R-MAT-24's degree 257..4096 bins run 1-2% faster (with the
// 32-bit counters, 7 still leaves the run 0.2-0.3% slower).  At p <= 6 one-copy teams run at 7 (no spill, where 8
// spilled 7 VGPRs): ER (n = 65,536, p = 6) 44.7 -> 43.5 us per launch; the k-mer graph's
// dense iterations run k_union_rows instead.
template <int P, int T>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(P >= 7 ? (T >= 32 ? 6 : 8) : (T == (1 << P) / 16 ? 7 : 8)))) void k_union(
    HbDev d, const uint32_t *__restrict__ list, uint32_t base, uint32_t count, int slot_id)
{
    const HbIterState *st = d.st;
    if (st->done) return;
    if constexpr (T == (1 << P) / 16 && (P == 5 || P == 6))
        if (d.rows && !list && (base & 63u) == 0 && rows_iteration(d, *st)) return;  // k_union_rows runs it
    HB_TABLES_LDS(P, d, tb)
    const uint32_t it = st->iter;
    if (st->push) {  // only the bin's nodes with a changed neighbour (k_push_compact)
        list = d.dyn_list + d.dyn_base[slot_id];
        count = d.dyn_count[slot_id];
    }
    HbView v;
    v.off = d.off;
    v.nbr = d.nbr;
    v.orig = d.orig;
    v.node_id = d.node_id;
    v.in = st->parity ? d.regs1 : d.regs0;
    v.out = st->parity ? d.regs0 : d.regs1;
    v.est = d.est;
    v.dist = d.dist;
    v.inv = d.inv;
    v.bitsR = (it & 1u) ? d.bits0 : d.bits1;
    v.bitsW = (it & 1u) ? d.bits1 : d.bits0;
    v.t = st->t;
    v.inv_t = st->inv_t;
    v.use_frontier = d.frontier && it > 1 && !st->skip_scan && !st->push;
    // outside dense iterations only the neighbours that changed at t-1 are gathered
    v.gather_bits = d.frontier && it > 1 && !st->skip_scan ? v.bitsR : nullptr;
    __shared__ __attribute__((aligned(16))) uint32_t s_first[kBlock * 16];
    v.first = it == 1 && P <= 10;
    v.scr = s_first;
    const uint32_t tl = threadIdx.x % T;
    const uint32_t teams_per_grid = gridDim.x * (kBlock / T);
    HbCounts c;
    if constexpr (T == (1 << P) / 16 && P != 5 && P != 6) {
        // dense iteration over an id-range bin: each wavefront's teams hold consecutive
        // nodes, and neighbours among them come from the wavefront's LDS window (at p = 5
        // and 6 k_union_rows runs these iterations)
        if (d.window && !list && !v.first && !v.gather_bits) {
            uint4 *win = reinterpret_cast<uint4 *>(s_first) + (threadIdx.x / HB_WAVE) * HB_WAVE;
            const uint32_t in_wave = (threadIdx.x % HB_WAVE) / T;
            for (uint32_t team = blockIdx.x * (kBlock / T) + threadIdx.x / T; team < count; team += teams_per_grid) {
                const uint32_t lo = base + team - in_wave;
                union_node_window<P, T>(v, tb, base + team, tl, c, lo, min(lo + HB_WAVE / T, base + count), win);
            }
            block_add5(c.inc, c.arcs, c.nodes, c.chg, c.gathered, d.rec + it % HB_REC_RING, slot_id);
            return;
        }
    }
    if (v.gather_bits) {
        for (uint32_t team = blockIdx.x * (kBlock / T) + threadIdx.x / T; team < count; team += teams_per_grid)
            union_node<P, T, true>(v, tb, list ? list[team] : base + team, tl, c);
    } else {
        for (uint32_t team = blockIdx.x * (kBlock / T) + threadIdx.x / T; team < count; team += teams_per_grid)
            union_node<P, T, false>(v, tb, list ? list[team] : base + team, tl, c);
    }
    block_add5(c.inc, c.arcs, c.nodes, c.chg, c.gathered, d.rec + it % HB_REC_RING, slot_id);
}

// ---------------------------------------------------------------------------
// hubs, phase 1: one wavefront per chunk of a hub's neighbour list writes the
// chunk's partial union to scratch (not the hub's own registers)
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(kBlock) void k_chunk(HbDev d, const uint32_t *__restrict__ tnode,
                                                  const uint32_t *__restrict__ tbeg,
                                                  const uint32_t *__restrict__ tend, uint32_t ntasks)
{
    constexpr uint32_t M = 1u << P, G = M / 16, T = HB_WAVE, S = T / G;
    const HbIterState *st = d.st;
    if (st->done) return;
    const uint32_t it = st->iter;
    const uint8_t *__restrict__ in = st->parity ? d.regs1 : d.regs0;
    const uint32_t *__restrict__ bitsR = (it & 1u) ? d.bits0 : d.bits1;
    const bool use_frontier = d.frontier && it > 1 && !st->skip_scan;
    // push iteration: a hub whose active bit is clear has no changed neighbour at all
    const bool push = st->push != 0;
    const bool first = it == 1;
    __shared__ uint32_t s_first[kBlock * 16];
    const uint32_t tl = threadIdx.x % T, cl = tl % G, slot = tl / G;
    const uint32_t waves_per_grid = gridDim.x * (kBlock / T);
    unsigned long long c_arcs = 0, c_gathered = 0;

    for (uint32_t task = blockIdx.x * (kBlock / T) + threadIdx.x / T; task < ntasks;
         task += waves_per_grid) {
        const uint32_t eb = tbeg[task], ee = tend[task];
        const uint32_t *__restrict__ nb = d.nbr + eb;
        const uint32_t *__restrict__ nbh = d.orig + eb;
        const uint32_t len = ee - eb;
        bool active = !(push && !bit_test(hb_active(d, it), tnode[task]));
        if (use_frontier && active) {
            uint32_t any = 0;
            for (uint32_t j = tl; j < len; j += T) any |= bit_test(bitsR, nb[j]);
            active = hb_copy_or<T>(any) != 0;
        }
        if (tl == 0) d.chunk_flag[task] = active ? 1u : 0u;
        if (!active) continue;
        uint32_t gathered = 0;  // by this lane's copy
        uint4 acc = use_frontier
                        ? hb_gather_union<P, S, true>(in, first, s_first, bitsR, nb, nbh, len, slot, cl,
                                                      make_uint4(0, 0, 0, 0), gathered)
                        : hb_gather_union<P, S, false>(in, first, s_first, nullptr, nb, nbh, len, slot, cl,
                                                       make_uint4(0, 0, 0, 0), gathered);
        if (cl == 0) c_gathered += gathered;
#pragma unroll
        for (uint32_t s = G; s < T; s <<= 1) acc = hb_bmax4(acc, hb_shfl_xor4<T>(acc, s));
        if (slot == 0) *(uint4 *)(d.scratch + ((uint64_t)task << P) + cl * 16) = acc;
        if (tl == 0) c_arcs += len;
    }
    block_add5(0, c_arcs, 0, 0, c_gathered, d.rec + it % HB_REC_RING, HB_SLOT_CHUNK);
}

// hubs, phase 2: one wavefront per hub merges its own registers with the
// partials of its recomputed chunks (a chunk that was not recomputed has no
// changed neighbour, so its union is already contained in c_{t-1}(u))
template <int P>
__global__ __launch_bounds__(kBlock) void k_merge(HbDev d, const uint32_t *__restrict__ hnode,
                                                  const uint32_t *__restrict__ htask0,
                                                  const uint32_t *__restrict__ hntask, uint32_t nhubs)
{
    constexpr uint32_t M = 1u << P, G = M / 16, T = HB_WAVE, S = T / G;
    const HbIterState *st = d.st;
    if (st->done) return;
    HB_TABLES_LDS(P, d, tb)
    const uint32_t it = st->iter;
    const double t = st->t, inv_t = st->inv_t;
    const uint8_t *__restrict__ in = st->parity ? d.regs1 : d.regs0;
    uint8_t *__restrict__ out = st->parity ? d.regs0 : d.regs1;
    const uint32_t *__restrict__ bitsR = (it & 1u) ? d.bits0 : d.bits1;
    uint32_t *bitsW = (it & 1u) ? d.bits1 : d.bits0;
    const uint32_t tl = threadIdx.x % T, cl = tl % G, slot = tl / G;
    const uint32_t waves_per_grid = gridDim.x * (kBlock / T);
    unsigned long long c_inc = 0, c_nodes = 0, c_chg = 0;

    for (uint32_t h = blockIdx.x * (kBlock / T) + threadIdx.x / T; h < nhubs; h += waves_per_grid) {
        const uint32_t u = hnode[h], t0 = htask0[h], nt = hntask[h];
        uint32_t any = 0;
        for (uint32_t j = tl; j < nt; j += T) any |= d.chunk_flag[t0 + j];
        const bool active = hb_copy_or<T>(any) != 0;
        if (!active) {
            if (slot == 0 && bit_test(bitsR, u))
                *(uint4 *)(out + ((uint64_t)u << P) + cl * 16) =
                    *(const uint4 *)(in + ((uint64_t)u << P) + cl * 16);
            continue;
        }
        const uint4 own = *(const uint4 *)(in + ((uint64_t)u << P) + cl * 16);
        uint4 acc = own;
        for (uint32_t j = slot; j < nt; j += S)
            if (d.chunk_flag[t0 + j])
                acc = hb_bmax4(acc, *(const uint4 *)(d.scratch + ((uint64_t)(t0 + j) << P) + cl * 16));
#pragma unroll
        for (uint32_t s = G; s < T; s <<= 1) acc = hb_bmax4(acc, hb_shfl_xor4<T>(acc, s));
        const uint32_t diffw = (acc.x != own.x) | (acc.y != own.y) | (acc.z != own.z) | (acc.w != own.w);
        const bool changed = hb_copy_or<G>(diffw) != 0;
        const float curr = hb_estimate_copy<G>(tb, acc);
        if (slot == 0) *(uint4 *)(out + ((uint64_t)u << P) + cl * 16) = acc;
        if (tl == 0) {
            const float prev = d.est[u];
            if (accumulate(d, u, prev, curr, t, inv_t)) c_inc += 1;
            d.est[u] = curr;
            if (changed) {
                atomicOr(bitsW + (u >> 5), 1u << (u & 31u));
                c_chg += 1;
            }
            c_nodes += 1;
        }
    }
    block_add5(c_inc, 0, c_nodes, c_chg, 0, d.rec + it % HB_REC_RING, HB_SLOT_MERGE);
}

// ---------------------------------------------------------------------------
// precisions 11..16: one wavefront per node walks the register row in 1 KiB
// chunks (64 lanes x 16 registers); a per-wave LDS histogram of register values
// gives zeros, the maximum and the exact integer form of sum 2^-M; a row whose
// sum is not exact in f32 replays the sequential loop from L2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void hist_add_word(uint32_t *h, uint32_t x)
{
    atomicAdd(h + (x & 0xffu), 1u);
    atomicAdd(h + ((x >> 8) & 0xffu), 1u);
    atomicAdd(h + ((x >> 16) & 0xffu), 1u);
    atomicAdd(h + (x >> 24), 1u);
}

// estimate from a 64-bin histogram of the row (all lanes of the wave return it)
template <int P>
__device__ float estimate_wide(const HbDev &d, const uint32_t *h, const uint8_t *row)
{
    constexpr uint32_t M = 1u << P;
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1);
    const uint32_t z = h[0];
    if (z > 0) {
        const float lc = d.lc[z];
        if ((double)lc <= d.lc_thr) return lc;
    }
    const uint32_t hv = h[lane];
    uint32_t mx = hv ? lane : 0u;
    for (int s = 1; s < HB_WAVE; s <<= 1) mx = max(mx, (uint32_t)__shfl_xor(mx, s));
    // k = sum_v h[v] 2^(mx - v), saturated at 2^24
    uint32_t k = 0;
    if (hv) {
        const uint32_t sh = mx - lane;
        k = sh >= 24u ? (1u << 24) : min(hv << sh, 1u << 24);
        if (sh < 24u && (hv >> (24u - sh)) != 0u) k = 1u << 24;
    }
    for (int s = 1; s < HB_WAVE; s <<= 1) k = min(k + (uint32_t)__shfl_xor(k, s), 1u << 24);
    float S;
    if (k < (1u << 24)) {
        S = ldexpf((float)k, -(int)mx);
    } else {
        S = 0.0f;
        if (lane == 0) {
            __threadfence();
            for (uint32_t i = 0; i < M; i += 4) {
                const uint32_t w = __hip_atomic_load((const uint32_t *)(row + i), __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
                for (int b = 0; b < 4; ++b) S = __fadd_rn(S, hb_pow2neg((w >> (8 * b)) & 0xffu));
            }
        }
        S = __shfl(S, 0);
    }
    float est = (float)((double)d.am / (double)S);
    if (est <= d.five_m) est = (float)((double)est - hb_bias_lookup(d.bias, est));
    return est;
}

template <int P>
__global__ __launch_bounds__(kBlock) void k_init_wide(HbDev d)
{
    constexpr uint32_t M = 1u << P;
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t u = (uint32_t)(gtid / HB_WAVE), lane = (uint32_t)(gtid % HB_WAVE);
    if (u >= d.n) return;
    uint32_t idx, rank;
    hb_index_rank(d.node_id ? d.node_id[u] : u, P, &idx, &rank);
    const bool isolated = d.off[u + 1] == d.off[u];
    for (uint32_t c = 0; c < M / 1024; ++c) {
        const uint32_t base = c * 1024 + lane * 16;
        uint32_t w[4] = {0, 0, 0, 0};
        if (idx >= base && idx < base + 16) w[(idx - base) / 4] = rank << (8 * (idx % 4));
        const uint4 r = make_uint4(w[0], w[1], w[2], w[3]);
        *(uint4 *)(d.regs0 + ((uint64_t)u << P) + base) = r;
        if (isolated) *(uint4 *)(d.regs1 + ((uint64_t)u << P) + base) = r;
    }
    if (lane == 0) {
        d.est[u] = d.lc[M - 1];  // one non-zero register: linear counting with m-1 zeros
        d.dist[u] = 0.0;
        d.inv[u] = 0.0;
    }
}

template <int P>
__global__ __launch_bounds__(kBlock) void k_union_wide(HbDev d, const uint32_t *__restrict__ list,
                                                       uint32_t count, int slot_id)
{
    constexpr uint32_t M = 1u << P;
    const HbIterState *st = d.st;
    if (st->done) return;
    __shared__ uint32_t hist[kBlock / HB_WAVE][64];
    const uint32_t it = st->iter;
    const double t = st->t, inv_t = st->inv_t;
    const uint8_t *__restrict__ in = st->parity ? d.regs1 : d.regs0;
    uint8_t *out = st->parity ? d.regs0 : d.regs1;
    const uint32_t *__restrict__ bitsR = (it & 1u) ? d.bits0 : d.bits1;
    uint32_t *bitsW = (it & 1u) ? d.bits1 : d.bits0;
    const bool use_frontier = d.frontier && it > 1 && !st->skip_scan;
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1), w = threadIdx.x / HB_WAVE;
    uint32_t *h = hist[w];
    unsigned long long c_inc = 0, c_arcs = 0, c_nodes = 0, c_chg = 0;

    for (uint32_t i = blockIdx.x * (kBlock / HB_WAVE) + w; i < count; i += gridDim.x * (kBlock / HB_WAVE)) {
        const uint32_t u = list[i];
        const uint32_t e0 = d.off[u], deg = d.off[u + 1] - e0;
        const uint32_t *__restrict__ nb = d.nbr + e0;
        bool active = true;
        if (use_frontier) {
            uint32_t any = 0;
            for (uint32_t j = lane; j < deg; j += HB_WAVE) any |= bit_test(bitsR, nb[j]);
            active = hb_copy_or<HB_WAVE>(any) != 0;
        }
        const uint8_t *own_row = in + ((uint64_t)u << P);
        uint8_t *out_row = out + ((uint64_t)u << P);
        if (!active) {
            if (bit_test(bitsR, u))
                for (uint32_t c = 0; c < M / 1024; ++c)
                    *(uint4 *)(out_row + c * 1024 + lane * 16) = *(const uint4 *)(own_row + c * 1024 + lane * 16);
            continue;
        }
        h[lane] = 0u;
        uint32_t diffw = 0;
        for (uint32_t c = 0; c < M / 1024; ++c) {
            const uint32_t off16 = c * 1024 + lane * 16;
            const uint4 own = *(const uint4 *)(own_row + off16);
            HbAcc pa = hb_acc(own);
            for (uint32_t j = 0; j < deg; ++j) hb_acc_max(pa, *(const uint4 *)(in + ((uint64_t)nb[j] << P) + off16));
            const uint4 acc = hb_acc_get(pa);
            diffw |= (acc.x != own.x) | (acc.y != own.y) | (acc.z != own.z) | (acc.w != own.w);
            *(uint4 *)(out_row + off16) = acc;
            hist_add_word(h, acc.x);
            hist_add_word(h, acc.y);
            hist_add_word(h, acc.z);
            hist_add_word(h, acc.w);
        }
        const bool changed = hb_copy_or<HB_WAVE>(diffw) != 0;
        const float curr = estimate_wide<P>(d, h, out_row);
        if (lane == 0) {
            const float prev = d.est[u];
            if (accumulate(d, u, prev, curr, t, inv_t)) c_inc += 1;
            d.est[u] = curr;
            if (changed) {
                atomicOr(bitsW + (u >> 5), 1u << (u & 31u));
                c_chg += 1;
            }
            c_arcs += deg;
            c_nodes += 1;
        }
    }
    block_add5(c_inc, c_arcs, c_nodes, c_chg, c_arcs, d.rec + it % HB_REC_RING, slot_id);  // gathers every row
}

template <int P>
__global__ __launch_bounds__(kBlock) void k_estimates_wide(HbDev d, const uint8_t *__restrict__ regs,
                                                           float *__restrict__ out)
{
    constexpr uint32_t M = 1u << P;
    __shared__ uint32_t hist[kBlock / HB_WAVE][64];
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1), w = threadIdx.x / HB_WAVE;
    uint32_t *h = hist[w];
    for (uint32_t u = blockIdx.x * (kBlock / HB_WAVE) + w; u < d.n; u += gridDim.x * (kBlock / HB_WAVE)) {
        const uint8_t *row = regs + ((uint64_t)u << P);
        h[lane] = 0u;
        for (uint32_t c = 0; c < M / 1024; ++c) {
            const uint4 r = *(const uint4 *)(row + c * 1024 + lane * 16);
            hist_add_word(h, r.x);
            hist_add_word(h, r.y);
            hist_add_word(h, r.z);
            hist_add_word(h, r.w);
        }
        const float e = estimate_wide<P>(d, h, row);
        if (lane == 0) out[u] = e;
    }
}

// ---------------------------------------------------------------------------
// neighbourhood function N(t) = sum_u est_t(u), summed exactly in 128-bit fixed
// point with 40 fractional bits (every estimate is >= 1, whose f32 ulp >= 2^-23)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_nf(HbDev d)
{
    const HbIterState *st = d.st;
    if (st->done) return;
    unsigned long long lo = 0, hi = 0;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < d.n; u += gridDim.x * blockDim.x) {
        const uint32_t b = __float_as_uint(d.est[u]);
        const int e = (int)((b >> 23) & 0xffu);
        const unsigned long long mant = (e ? (1ull << 23) : 0ull) | (b & 0x7fffffu);
        const int sh = (e ? e : 1) - 127 - 23 + 40;  // value * 2^40 = mant * 2^sh
        unsigned long long vlo, vhi;
        if (sh >= 64) { vlo = 0; vhi = mant << (sh - 64); }
        else if (sh > 0) { vlo = mant << sh; vhi = sh > 40 ? (mant >> (64 - sh)) : 0ull; }
        else { vlo = mant >> (-sh); vhi = 0; }
        const unsigned long long nlo = lo + vlo;
        hi += vhi + (nlo < lo);
        lo = nlo;
    }
    // wave reduction with carries
#pragma unroll
    for (int s = 1; s < HB_WAVE; s <<= 1) {
        const unsigned long long olo = __shfl_xor(lo, s), ohi = __shfl_xor(hi, s);
        const unsigned long long nlo = lo + olo;
        hi = hi + ohi + (nlo < lo);
        lo = nlo;
    }
    if ((threadIdx.x & (HB_WAVE - 1)) == 0) {
        HbIterRec *r = d.rec + st->iter % HB_REC_RING;
        const unsigned long long old = atomicAdd(&r->nf_lo, lo);
        atomicAdd(&r->nf_hi, hi + ((old + lo) < old ? 1ull : 0ull));
    }
}

// estimates after iteration t <= snap_k (surrogate inputs)
__global__ __launch_bounds__(kBlock) void k_snapshot(HbDev d)
{
    const HbIterState *st = d.st;
    if (st->done || st->iter > d.snap_k) return;
    float *dst = d.snaps + (uint64_t)(st->iter - 1) * d.n;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < d.n; u += gridDim.x * blockDim.x) dst[u] = d.est[u];
}

// end of iteration t: termination rule (hyper_ball.rs:833-849) and buffer swap
__global__ void k_finalize(HbDev d)
{
    HbIterState *st = d.st;
    if (st->done || threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t it = st->iter;
    unsigned long long inc = 0, chg = 0;
    for (int k = 0; k < HB_NSLOT; ++k) {
        inc += d.rec[it % HB_REC_RING].inc[k];
        chg += d.rec[it % HB_REC_RING].chg[k];
    }
    if (inc == 0 || (st->max_iters && it >= st->max_iters)) {
        st->done = 1;
        st->executed = it;
        st->converged = inc == 0;
        return;
    }
    st->iter = it + 1;
    st->parity ^= 1u;
    st->skip_scan = 2 * chg > (unsigned long long)d.n_linked;
    // push when few nodes changed (on R-MAT-24 a push pass at 24% changed nodes whose lists
    // hold 2% of the arcs measured slower than the scan with selective gathers)
    st->push = d.sym && d.p <= 10 && d.frontier && 16 * chg <= (unsigned long long)d.n_linked;
    st->t = st->t + 1.0;
    st->inv_t = __ddiv_rn(1.0, st->t);
}

// ---------------------------------------------------------------------------
// epilogues
// ---------------------------------------------------------------------------
template <int P>
__global__ __launch_bounds__(kBlock) void k_estimates(HbDev d, const uint8_t *__restrict__ regs,
                                                      float *__restrict__ out)
{
    constexpr uint32_t M = 1u << P, G = M / 16;
    HB_TABLES_LDS(P, d, tb)
    const uint64_t gtid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t u = (uint32_t)(gtid / G), cl = (uint32_t)(gtid % G);
    if (u >= d.n) return;
    const uint4 r = *(const uint4 *)(regs + ((uint64_t)u << P) + cl * 16);
    const float e = hb_estimate_copy<G>(tb, r);
    if (cl == 0) out[u] = e;
}

// hyper_ball.rs:434-555: closeness (:447-466), harmonic (:500-518, guard on dist),
// Lin (:540-547, 1 when dist is not normal)
__global__ __launch_bounds__(kBlock) void k_centrality(HbDev d, const float *__restrict__ est,
                                                       int kind, int normalize, double *__restrict__ out)
{
    const double nf = (double)d.n - 1.0;
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < d.n; u += gridDim.x * blockDim.x) {
        const double dd = d.dist[u];
        const double e = (double)est[u];
        const bool normal = isnormal(dd);
        double r;
        if (kind == 0) r = !normal ? 0.0 : (normalize ? __ddiv_rn(__dsub_rn(e, 1.0), dd) : __ddiv_rn(nf, dd));
        else if (kind == 1) r = !normal ? 0.0 : (normalize ? __ddiv_rn(d.inv[u], __dsub_rn(e, 1.0)) : __ddiv_rn(d.inv[u], nf));
        else {
            const double x = __dsub_rn(e, 1.0);
            r = !normal ? 1.0 : __ddiv_rn(__dmul_rn(x, x), dd);
        }
        out[u] = r;
    }
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
static unsigned grid_for(uint64_t threads, unsigned cap)
{
    const uint64_t b = (threads + kBlock - 1) / kBlock;
    return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

template <int P>
static hipError_t launch_union_T(const HbDev &d, const uint32_t *list, uint32_t base, uint32_t count, int T,
                                 unsigned max_blocks, int slot, hipStream_t s)
{
    constexpr int G = (1 << P) / 16;
    const unsigned grid = grid_for((uint64_t)count * T, max_blocks);
    switch (T) {
#define HB_CASE(TT)                                                                        \
    case TT:                                                                               \
        if constexpr (TT >= G) {                                                           \
            k_union<P, TT><<<grid, kBlock, 0, s>>>(d, list, base, count, slot);            \
            return hipGetLastError();                                                      \
        }                                                                                  \
        break;
        HB_CASE(1) HB_CASE(2) HB_CASE(4) HB_CASE(8) HB_CASE(16) HB_CASE(32) HB_CASE(64)
#undef HB_CASE
    }
    return hipErrorInvalidValue;
}

#define HB_DISPATCH_P(p, ...)                              \
    switch (p) {                                           \
    case 4: { constexpr int P = 4; __VA_ARGS__; } break;   \
    case 5: { constexpr int P = 5; __VA_ARGS__; } break;   \
    case 6: { constexpr int P = 6; __VA_ARGS__; } break;   \
    case 7: { constexpr int P = 7; __VA_ARGS__; } break;   \
    case 8: { constexpr int P = 8; __VA_ARGS__; } break;   \
    case 9: { constexpr int P = 9; __VA_ARGS__; } break;   \
    case 10: { constexpr int P = 10; __VA_ARGS__; } break; \
    default: return hipErrorInvalidValue;                  \
    }

#define HB_DISPATCH_WIDE(p, ...)                            \
    switch (p) {                                            \
    case 11: { constexpr int P = 11; __VA_ARGS__; } break;  \
    case 12: { constexpr int P = 12; __VA_ARGS__; } break;  \
    case 13: { constexpr int P = 13; __VA_ARGS__; } break;  \
    case 14: { constexpr int P = 14; __VA_ARGS__; } break;  \
    case 15: { constexpr int P = 15; __VA_ARGS__; } break;  \
    case 16: { constexpr int P = 16; __VA_ARGS__; } break;  \
    default: return hipErrorInvalidValue;                   \
    }

hipError_t launch_init(const HbDev &d, hipStream_t s)
{
    if (d.p > 10) {
        HB_DISPATCH_WIDE(d.p, {
            k_init_wide<P><<<(unsigned)(((uint64_t)d.n * HB_WAVE + kBlock - 1) / kBlock), kBlock, 0, s>>>(d);
        });
        return hipGetLastError();
    }
    // k_init stores the linear-counting estimate of m - 1 zeros for every node; it is what
    // the estimator returns for a one-insertion sketch when LC <= the HLL++ threshold,
    // which holds at every p (LC < 1.04, thresholds >= 10), checked here with the same
    // expression the LC table is filled with
    const double m = (double)(1u << d.p);
    if ((double)(float)(m * std::log(m / (m - 1.0))) > d.lc_thr) return hipErrorInvalidValue;
    HB_DISPATCH_P(d.p, {
        k_init<P><<<(unsigned)(((uint64_t)d.n + kBlock - 1) / kBlock), kBlock, 0, s>>>(d);
    });
    return hipGetLastError();
}

hipError_t launch_fill(uint32_t *p, uint64_t count, uint32_t value, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    k_fill_u32<<<grid_for(count, 4096), kBlock, 0, s>>>(p, count, value);
    return hipGetLastError();
}

hipError_t launch_iteration(const HbDev &d, const HbSchedule &sc, bool nf, hipStream_t s,
                            hipEvent_t *ev)
{
    const uint32_t words = (d.n + 31) / 32;
    const bool push_possible = d.sym && d.p <= 10 && d.frontier;
    // where push iterations can occur k_begin also marks, one wavefront per bitmap word
    k_begin<<<push_possible ? grid_for((uint64_t)words * HB_WAVE, 8192) : grid_for(words, 2048), kBlock, 0, s>>>(
        d, words);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (push_possible) {  // returns at once unless this is a push iteration
        HB_DISPATCH_P(d.p, k_push_compact<P><<<grid_for(words, 4096), kBlock, 0, s>>>(d, words));
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    for (int b = 0; b < sc.nbins; ++b) {
        if (sc.bin_count[b] == 0) continue;
        if (ev) (void)hipEventRecord(ev[2 * b], s);
        if (d.p > 10) {
            HB_DISPATCH_WIDE(d.p, {
                k_union_wide<P><<<grid_for((uint64_t)sc.bin_count[b] * HB_WAVE, sc.max_blocks), kBlock, 0, s>>>(
                    d, sc.bin_list[b], sc.bin_count[b], b);
                e = hipGetLastError();
            });
        } else {
            // k_union_rows and k_union each return at once in the iterations the other runs
            if (rows_bin(d, sc, b)) {
                const unsigned grid = grid_for(sc.bin_count[b], sc.max_blocks);
                if (d.p == 5) k_union_rows<5><<<grid, kBlock, 0, s>>>(d, sc.bin_base[b], sc.bin_count[b], b);
                else k_union_rows<6><<<grid, kBlock, 0, s>>>(d, sc.bin_base[b], sc.bin_count[b], b);
                e = hipGetLastError();
                if (e != hipSuccess) return e;
            }
            HB_DISPATCH_P(d.p, e = launch_union_T<P>(d, sc.bin_list[b], sc.bin_base[b], sc.bin_count[b],
                                                     sc.bin_team[b], sc.max_blocks, b, s));
        }
        if (ev) (void)hipEventRecord(ev[2 * b + 1], s);
        if (e != hipSuccess) return e;
    }
    if (sc.ntasks) {
        HB_DISPATCH_P(d.p, {
            if (ev) (void)hipEventRecord(ev[2 * HB_SLOT_CHUNK], s);
            k_chunk<P><<<grid_for((uint64_t)sc.ntasks * HB_WAVE, sc.max_blocks), kBlock, 0, s>>>(
                d, sc.task_node, sc.task_beg, sc.task_end, sc.ntasks);
            if (ev) (void)hipEventRecord(ev[2 * HB_SLOT_CHUNK + 1], s);
            if (ev) (void)hipEventRecord(ev[2 * HB_SLOT_MERGE], s);
            k_merge<P><<<grid_for((uint64_t)sc.nhubs * HB_WAVE, sc.max_blocks), kBlock, 0, s>>>(
                d, sc.hub_node, sc.hub_task0, sc.hub_ntask, sc.nhubs);
            if (ev) (void)hipEventRecord(ev[2 * HB_SLOT_MERGE + 1], s);
        });
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nf) {
        k_nf<<<grid_for(d.n, 1024), kBlock, 0, s>>>(d);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (d.snap_k) {
        k_snapshot<<<grid_for(d.n, 4096), kBlock, 0, s>>>(d);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    k_finalize<<<1, 64, 0, s>>>(d);
    return hipGetLastError();
}

hipError_t arc_fingerprints(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint64_t arcs,
                            unsigned long long *work, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    k_arc_fingerprint<<<kFingerprintBlocks, kBlock, 0, s>>>(off, nbr, n, arcs, work + 4);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return sum_columns(work + 4, kFingerprintBlocks, 4, work, s);
}

// out[k] = sum over r < rows of part[r * cols + k], one block per column, a fixed
// summation order (so f64 sums are reproducible)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sum_columns(const T *part, uint32_t rows, uint32_t cols, T *out)
{
    __shared__ T red[kBlock];
    const uint32_t k = blockIdx.x;
    T t = 0;
    for (uint32_t r = threadIdx.x; r < rows; r += kBlock) t += part[(uint64_t)r * cols + k];
    red[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t w = kBlock / 2; w; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
}

template <typename T>
hipError_t sum_columns(const T *part, uint32_t rows, uint32_t cols, T *out, hipStream_t s)
{
    if (cols == 0) return hipSuccess;
    k_sum_columns<T><<<cols, kBlock, 0, s>>>(part, rows, cols, out);
    return hipGetLastError();
}
template hipError_t sum_columns<unsigned long long>(const unsigned long long *, uint32_t, uint32_t,
                                                    unsigned long long *, hipStream_t);
template hipError_t sum_columns<double>(const double *, uint32_t, uint32_t, double *, hipStream_t);

hipError_t launch_estimates(const HbDev &d, const uint8_t *regs, float *out, hipStream_t s)
{
    if (d.p > 10) {
        HB_DISPATCH_WIDE(d.p, {
            k_estimates_wide<P><<<grid_for((uint64_t)d.n * HB_WAVE, 8192), kBlock, 0, s>>>(d, regs, out);
        });
        return hipGetLastError();
    }
    HB_DISPATCH_P(d.p, {
        const uint64_t threads = (uint64_t)d.n * ((1 << P) / 16);
        k_estimates<P><<<(unsigned)((threads + kBlock - 1) / kBlock), kBlock, 0, s>>>(d, regs, out);
    });
    return hipGetLastError();
}

hipError_t launch_centrality(const HbDev &d, const float *est, int kind, int normalize, double *out,
                             hipStream_t s)
{
    k_centrality<<<grid_for(d.n, 4096), kBlock, 0, s>>>(d, est, kind, normalize, out);
    return hipGetLastError();
}

}  // namespace hb
