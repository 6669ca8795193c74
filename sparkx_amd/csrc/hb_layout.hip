// hb_layout.hip -- degree-ordered device layout of a staged CSR, and the gathers that
// bring per-node results back to the caller's node order.
//
// HyperBall's union reads, for every arc, the 2^p-byte register block of a random
// neighbour; its speed is set by how many such random requests the memory system
// serves per second, which rises when the requests fall into a smaller address range.
// Renumbering the nodes by descending degree (ties by id) puts the most-gathered
// blocks at the start of the register array (on R-MAT-24, 79% of all gathers land in
// the first 64 MiB), makes every degree bin a contiguous id range (its own rows,
// estimates and accumulators are then read and written in order), and with each list
// sorted by the new ids a node walks its neighbours' blocks in address order.
//
// Results do not depend on the numbering: a sketch is the union of its neighbours'
// sketches, and every node still inserts its original id (rorig / perm), so registers,
// estimates, accumulators and iteration counts are those of the input order, moved.
#include <algorithm>
#include <cstdint>
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include "hb_kernels.h"

namespace hb {
namespace {

constexpr int kBlock = 256;

unsigned grid_for(uint64_t items, unsigned per_block = kBlock)
{
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + per_block - 1) / per_block, 16384));
}

__global__ __launch_bounds__(kBlock) void k_deg_ids(const uint32_t *off, uint32_t n, uint32_t *deg, uint32_t *ids)
{
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        deg[u] = off[u + 1] - off[u];
        ids[u] = u;
    }
}

__global__ __launch_bounds__(kBlock) void k_invert(const uint32_t *perm, uint32_t n, uint32_t *sigma)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) sigma[perm[i]] = i;
}

// one wavefront per new node copies its old list, renumbered
__global__ __launch_bounds__(kBlock) void k_fill_lists(const uint32_t *off, const uint32_t *nbr, const uint32_t *perm,
                                                       const uint32_t *sigma, const uint32_t *roff, uint32_t n,
                                                       uint32_t *rnbr, uint32_t *rorig)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (blockDim.x / 64u);
    for (uint32_t i = blockIdx.x * (blockDim.x / 64u) + threadIdx.x / 64u; i < n; i += waves) {
        const uint32_t u = perm[i];
        const uint32_t b = off[u], d = off[u + 1] - b, o = roff[i];
        for (uint32_t k = lane; k < d; k += 64u) {
            const uint32_t w = nbr[b + k];
            rnbr[o + k] = sigma[w];
            rorig[o + k] = w;
        }
    }
}

template <typename W>
__global__ __launch_bounds__(kBlock) void k_unpermute(const W *src, W *dst, const uint32_t *sigma, uint32_t n,
                                                      uint32_t words)
{
    const uint64_t total = (uint64_t)n * words;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t old = i / words, k = i - old * words;
        dst[i] = src[(uint64_t)sigma[old] * words + k];
    }
}

struct Tmp {
    void *p = nullptr;
    ~Tmp()
    {
        if (p) (void)hipFree(p);
    }
};

#define HBL_TRY(call)                     \
    do {                                  \
        const hipError_t e_ = (call);     \
        if (e_ != hipSuccess) return e_;  \
    } while (0)

}  // namespace

namespace {
// Given perm (new id -> caller's id) and degs[i] = degree of perm[i]: sigma, the
// renumbered offsets, and every list in new-node order renumbered and sorted by new
// id with the caller's ids carried along.  Synchronises s.
hipError_t finish_relabel(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint32_t arcs, const uint32_t *perm,
                          const uint32_t *degs, uint32_t *sigma, uint32_t *roff, uint32_t *rnbr, uint32_t *rorig,
                          hipStream_t s)
{
    k_invert<<<grid_for(n), kBlock, 0, s>>>(perm, n, sigma);
    HBL_TRY(hipGetLastError());
    HBL_TRY(hipMemsetAsync(roff, 0, 4, s));
    size_t scan_bytes = 0;
    HBL_TRY(rocprim::inclusive_scan(nullptr, scan_bytes, degs, roff + 1, (size_t)n, rocprim::plus<uint32_t>(), s));
    Tmp scan_tmp;
    HBL_TRY(hipMalloc(&scan_tmp.p, std::max<size_t>(scan_bytes, 1)));
    HBL_TRY(rocprim::inclusive_scan(scan_tmp.p, scan_bytes, degs, roff + 1, (size_t)n, rocprim::plus<uint32_t>(), s));
    if (arcs == 0) return hipStreamSynchronize(s);
    // lists in new node order, renumbered, then each sorted by new id (carrying the old id)
    Tmp kin, vin;
    HBL_TRY(hipMalloc(&kin.p, (size_t)arcs * 4));
    HBL_TRY(hipMalloc(&vin.p, (size_t)arcs * 4));
    k_fill_lists<<<grid_for((uint64_t)n * 64), kBlock, 0, s>>>(off, nbr, perm, sigma, roff, n, (uint32_t *)kin.p,
                                                                (uint32_t *)vin.p);
    HBL_TRY(hipGetLastError());
    const unsigned bits = std::max(1u, 32u - (unsigned)__builtin_clz(std::max(n - 1, 1u)));
    size_t seg_bytes = 0;
    HBL_TRY(rocprim::segmented_radix_sort_pairs(nullptr, seg_bytes, (uint32_t *)kin.p, rnbr, (uint32_t *)vin.p,
                                                rorig, arcs, n, roff, roff + 1, 0, bits, s));
    Tmp seg_tmp;
    HBL_TRY(hipMalloc(&seg_tmp.p, std::max<size_t>(seg_bytes, 1)));
    HBL_TRY(rocprim::segmented_radix_sort_pairs(seg_tmp.p, seg_bytes, (uint32_t *)kin.p, rnbr, (uint32_t *)vin.p,
                                                rorig, arcs, n, roff, roff + 1, 0, bits, s));
    return hipStreamSynchronize(s);  // the temporaries are freed on return
}

// ---- chain layout -------------------------------------------------------------
// Multi-source breadth-first search from a sparse pseudo-random set of seeds (about
// one linked node in kChainSeed); every node takes the (level, seed) of the nearest
// seed (the smaller seed on a tie), and nodes are ordered by (seed, level, id).  On a
// graph made of long chains (a de Bruijn graph of a genome) the nodes within a
// seed's reach are numbered along the chain, two per level (one each side), so the
// neighbours of a node sit a few rows away and share its cache lines.
constexpr uint32_t kChainSeed = 64;
constexpr uint64_t kKeyUnset = ~0ull - 1, kKeyIsolated = ~0ull;

__device__ __forceinline__ uint32_t chain_mix(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    return x ^ (x >> 16);
}

__global__ __launch_bounds__(kBlock) void k_chain_seed(const uint32_t *off, uint32_t n, unsigned long long *key,
                                                       uint32_t *frontier, uint32_t *count)
{
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const bool linked = off[u + 1] != off[u];
        const bool seed = linked && chain_mix(u) % kChainSeed == 0;
        key[u] = !linked ? kKeyIsolated : seed ? (unsigned long long)u : kKeyUnset;  // level 0 | seed u
        if (seed) frontier[atomicAdd(count, 1u)] = u;
    }
}

// one level: every frontier node offers (level + 1, its seed) to its unreached
// neighbours; the first offer to reach a node also lists it in the next frontier
__global__ __launch_bounds__(kBlock) void k_chain_expand(const uint32_t *off, const uint32_t *nbr, unsigned long long *key,
                                                         const uint32_t *frontier, const uint32_t *fcount,
                                                         uint32_t *next, uint32_t *ncount)
{
    const uint32_t fc = *fcount;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < fc; i += gridDim.x * blockDim.x) {
        const uint32_t u = frontier[i];
        const unsigned long long ku = key[u];
        const unsigned long long offer = ku + (1ull << 32);  // level + 1, same seed
        for (uint32_t e = off[u]; e < off[u + 1]; ++e) {
            const uint32_t v = nbr[e];
            if (key[v] != kKeyUnset) continue;
            if (atomicMin(key + v, offer) == kKeyUnset) next[atomicAdd(ncount, 1u)] = v;
        }
    }
}

// sort key (seed, level) for reached nodes; unreached linked nodes, then isolated ones
__global__ __launch_bounds__(kBlock) void k_chain_sortkey(unsigned long long *key, uint32_t n, uint32_t *ids)
{
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const unsigned long long k = key[u];
        key[u] = k == kKeyIsolated ? 0xffffffff00000000ull
                 : k == kKeyUnset  ? 0xfffffffe00000000ull
                                   : ((k & 0xffffffffull) << 32) | (k >> 32);
        ids[u] = u;
    }
}

__global__ __launch_bounds__(kBlock) void k_perm_deg(const uint32_t *off, const uint32_t *perm, uint32_t n,
                                                     uint32_t *degs)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        degs[i] = off[perm[i] + 1] - off[perm[i]];
}
}  // namespace

hipError_t relabel_by_degree(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint32_t arcs, uint32_t *perm,
                             uint32_t *sigma, uint32_t *roff, uint32_t *rnbr, uint32_t *rorig, hipStream_t s)
{
    if (n == 0) return hipMemsetAsync(roff, 0, 4, s);
    Tmp deg, ids, degs, tmp;
    HBL_TRY(hipMalloc(&deg.p, (size_t)n * 4));
    HBL_TRY(hipMalloc(&ids.p, (size_t)n * 4));
    HBL_TRY(hipMalloc(&degs.p, (size_t)n * 4));
    k_deg_ids<<<grid_for(n), kBlock, 0, s>>>(off, n, (uint32_t *)deg.p, (uint32_t *)ids.p);
    HBL_TRY(hipGetLastError());
    // stable: equal degrees keep ascending ids
    size_t bytes = 0;
    HBL_TRY(rocprim::radix_sort_pairs_desc(nullptr, bytes, (uint32_t *)deg.p, (uint32_t *)degs.p,
                                           (uint32_t *)ids.p, perm, n, 0, 32, s));
    HBL_TRY(hipMalloc(&tmp.p, std::max<size_t>(bytes, 1)));
    HBL_TRY(rocprim::radix_sort_pairs_desc(tmp.p, bytes, (uint32_t *)deg.p, (uint32_t *)degs.p, (uint32_t *)ids.p,
                                           perm, n, 0, 32, s));
    return finish_relabel(off, nbr, n, arcs, perm, (const uint32_t *)degs.p, sigma, roff, rnbr, rorig, s);
}

hipError_t relabel_by_chain(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint32_t arcs, uint32_t *perm,
                            uint32_t *sigma, uint32_t *roff, uint32_t *rnbr, uint32_t *rorig, uint32_t max_levels,
                            uint32_t *levels, hipStream_t s)
{
    *levels = 0;
    if (n == 0) return hipMemsetAsync(roff, 0, 4, s);
    Tmp key, keys, ids, fr0, fr1, counts, degs, tmp;
    HBL_TRY(hipMalloc(&key.p, (size_t)n * 8));
    HBL_TRY(hipMalloc(&keys.p, (size_t)n * 8));
    HBL_TRY(hipMalloc(&ids.p, (size_t)n * 4));
    HBL_TRY(hipMalloc(&fr0.p, (size_t)n * 4));
    HBL_TRY(hipMalloc(&fr1.p, (size_t)n * 4));
    HBL_TRY(hipMalloc(&counts.p, 8));
    uint32_t *cnt = (uint32_t *)counts.p;
    HBL_TRY(hipMemsetAsync(cnt, 0, 8, s));
    k_chain_seed<<<grid_for(n), kBlock, 0, s>>>(off, n, (unsigned long long *)key.p, (uint32_t *)fr0.p, cnt);
    HBL_TRY(hipGetLastError());
    uint32_t *fr[2] = {(uint32_t *)fr0.p, (uint32_t *)fr1.p};
    for (uint32_t level = 0; level < max_levels; ++level) {
        const uint32_t cur = level & 1u;
        uint32_t h = 0;
        HBL_TRY(hipMemcpyAsync(&h, cnt + cur, 4, hipMemcpyDeviceToHost, s));
        HBL_TRY(hipStreamSynchronize(s));
        if (h == 0) break;
        *levels = level + 1;
        HBL_TRY(hipMemsetAsync(cnt + (cur ^ 1u), 0, 4, s));
        k_chain_expand<<<grid_for(h), kBlock, 0, s>>>(off, nbr, (unsigned long long *)key.p, fr[cur], cnt + cur,
                                                      fr[cur ^ 1u], cnt + (cur ^ 1u));
        HBL_TRY(hipGetLastError());
    }
    k_chain_sortkey<<<grid_for(n), kBlock, 0, s>>>((unsigned long long *)key.p, n, (uint32_t *)ids.p);
    HBL_TRY(hipGetLastError());
    // stable: equal keys (unreached and isolated nodes) keep ascending ids
    size_t bytes = 0;
    HBL_TRY(rocprim::radix_sort_pairs(nullptr, bytes, (unsigned long long *)key.p, (unsigned long long *)keys.p,
                                      (uint32_t *)ids.p, perm, n, 0, 64, s));
    HBL_TRY(hipMalloc(&tmp.p, std::max<size_t>(bytes, 1)));
    HBL_TRY(rocprim::radix_sort_pairs(tmp.p, bytes, (unsigned long long *)key.p, (unsigned long long *)keys.p,
                                      (uint32_t *)ids.p, perm, n, 0, 64, s));
    HBL_TRY(hipMalloc(&degs.p, (size_t)n * 4));
    k_perm_deg<<<grid_for(n), kBlock, 0, s>>>(off, perm, n, (uint32_t *)degs.p);
    HBL_TRY(hipGetLastError());
    return finish_relabel(off, nbr, n, arcs, perm, (const uint32_t *)degs.p, sigma, roff, rnbr, rorig, s);
}

hipError_t unpermute_rows(const void *src, void *dst, const uint32_t *sigma, uint32_t n, size_t row_bytes,
                          hipStream_t s)
{
    if (n == 0 || row_bytes == 0) return hipSuccess;
    if (row_bytes % 16 == 0) {
        const uint32_t w = (uint32_t)(row_bytes / 16);
        k_unpermute<uint4><<<grid_for((uint64_t)n * w), kBlock, 0, s>>>((const uint4 *)src, (uint4 *)dst, sigma, n, w);
    } else if (row_bytes % 8 == 0) {
        const uint32_t w = (uint32_t)(row_bytes / 8);
        k_unpermute<uint64_t><<<grid_for((uint64_t)n * w), kBlock, 0, s>>>((const uint64_t *)src, (uint64_t *)dst,
                                                                            sigma, n, w);
    } else if (row_bytes % 4 == 0) {
        const uint32_t w = (uint32_t)(row_bytes / 4);
        k_unpermute<uint32_t><<<grid_for((uint64_t)n * w), kBlock, 0, s>>>((const uint32_t *)src, (uint32_t *)dst,
                                                                            sigma, n, w);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace hb
