// hb_exact.hip -- exact BFS centralities on the GPU: 64 breadth-first searches at
// once, one bit lane per source in a u64 per node, advanced one level per launch.
//
// Semantics follow the reference's BFS (src/trails/bfs.rs:199-234): the source is
// queued with distance 0 but not marked visited, so it is rediscovered one step
// after the first node with an arc back to it (distance 1 through a self-loop) and
// counted like any other node.  Level-synchronous form, per source bit:
//     visited = {}, frontier = {source}
//     level d: next = nodes with an arc from a frontier node;  newly = next \ visited
//              visited |= newly, frontier = newly, reachable += |newly|, total += d |newly|
// which discovers exactly the nodes (and distances) of the reference's FIFO queue.
// Arcs are followed forwards (push), so directed graphs are searched as the
// reference searches them; on symmetric graphs a level with a large frontier pulls
// instead (each node not yet reached ORs its neighbours' frontier words), which
// computes the same `next` without atomics.
//
// Per-source sums, bit-identical to the reference's:
//   reachable : integer count (BFSDists::recheable)
//   total     : Σ d as an integer; the reference's f64 running sum of integers is
//               exact while it stays below 2^53 (16.7M nodes at distance 1000 is 2^34)
//   inv_sum   : Σ_v 1/d(v) over v = 0..n-1 in node order, 1/d a correctly rounded
//               f64 division and 0 for unreached nodes (exact.rs:386-396: an iterator
//               sum over get_distances()), replayed sequentially per source from a
//               u16 distance per (node, source) -- the order fixes the rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <vector>

#include "../../include/sparkx_hb.h"
#include "hb_kernels.h"

namespace hb {

constexpr int kBfsBlock = 256;

__global__ __launch_bounds__(kBfsBlock) void k_bfs_init(uint64_t *frontier, uint64_t *visited, uint64_t *next,
                                                        uint32_t n, uint32_t s0, uint32_t count)
{
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x) {
        const uint32_t k = u - s0;  // wraps for u < s0
        frontier[u] = (u >= s0 && k < count) ? (1ull << k) : 0ull;
        visited[u] = 0ull;
        next[u] = 0ull;
    }
}

template <typename T>
__global__ __launch_bounds__(kBfsBlock) void k_bfs_zero(T *p, uint64_t count)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// next[v] |= frontier[u] for every arc u -> v whose bits are not all visited at v.
// A wavefront takes 64 consecutive nodes at a time: a frontier node of degree <= 64
// is walked by its own lane, a larger one by the whole wavefront, so hubs do not
// serialise on one lane; nodes of degree > hub_deg are left to k_bfs_hubs, which
// spreads their lists over many wavefronts.
__device__ __forceinline__ void hb_bfs_push_arc(const uint64_t *__restrict__ visited, uint64_t *next, uint32_t v,
                                                uint64_t f)
{
    const uint64_t add = f & ~visited[v];
    if (add) atomicOr((unsigned long long *)&next[v], (unsigned long long)add);
}

__global__ __launch_bounds__(kBfsBlock) void k_bfs_push(const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                                                        const uint64_t *__restrict__ frontier,
                                                        const uint64_t *__restrict__ visited, uint64_t *next, uint32_t n,
                                                        uint32_t hub_deg)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (kBfsBlock / 64);
    for (uint32_t base = (blockIdx.x * (kBfsBlock / 64) + threadIdx.x / 64) * 64; base < n; base += waves * 64) {
        const uint32_t u = base + lane;  // uniform loop bound: every lane takes part in the ballot
        uint64_t f = 0;
        uint32_t e0 = 0, e1 = 0;
        if (u < n) {
            f = frontier[u];
            if (f) {
                e0 = off[u];
                e1 = off[u + 1];
            }
        }
        if (e1 - e0 <= 64)
            for (uint32_t e = e0; e < e1; ++e) hb_bfs_push_arc(visited, next, nbr[e], f);
        unsigned long long big = __ballot(e1 - e0 > 64 && e1 - e0 <= hub_deg);
        while (big) {
            const int i = __ffsll(big) - 1;
            big &= big - 1;
            const uint32_t b0 = (uint32_t)__shfl((int)e0, i), b1 = (uint32_t)__shfl((int)e1, i);
            const uint64_t fi = (uint64_t)__shfl((long long)f, i);
            for (uint32_t e = b0 + lane; e < b1; e += 64) hb_bfs_push_arc(visited, next, nbr[e], fi);
        }
    }
}

// Level 1: the frontier is the pass's sources s0..s0+count-1, whose lists lie side by
// side in nbr[off[s0], off[s0 + count]).  One thread per arc (its source found by a
// binary search over the count + 1 offsets in LDS), so a pass of hub sources spreads
// over the whole grid instead of one wavefront.  Nothing is visited yet at level 1.
__global__ __launch_bounds__(kBfsBlock) void k_bfs_push_sources(const uint32_t *__restrict__ off,
                                                                const uint32_t *__restrict__ nbr, uint64_t *next,
                                                                uint32_t s0, uint32_t count)
{
    __shared__ uint32_t so[65];
    if (threadIdx.x <= count) so[threadIdx.x] = off[s0 + threadIdx.x];
    __syncthreads();
    const uint32_t e0 = so[0], e1 = so[count];
    for (uint64_t e = (uint64_t)e0 + blockIdx.x * blockDim.x + threadIdx.x; e < e1; e += gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = count;  // so[lo] <= e < so[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (so[mid] <= e) lo = mid; else hi = mid;
        }
        atomicOr((unsigned long long *)&next[nbr[e]], 1ull << lo);
    }
}

// OR of frontier[nbr[e]] over e = b0 + lane, b0 + lane + 64, ... < b1 (this lane's
// share of a list walked by a whole wavefront), four arcs' loads in flight at a time
__device__ __forceinline__ uint64_t hb_bfs_or_range(const uint32_t *__restrict__ nbr,
                                                    const uint64_t *__restrict__ frontier, uint32_t b0, uint32_t b1,
                                                    uint32_t lane)
{
    uint64_t g = 0;
    uint64_t e = (uint64_t)b0 + lane;
    for (; e + 3 * 64 < b1; e += 4 * 64) {
        const uint32_t w0 = nbr[e], w1 = nbr[e + 64], w2 = nbr[e + 128], w3 = nbr[e + 192];
        g |= frontier[w0] | frontier[w1] | frontier[w2] | frontier[w3];
    }
    for (; e < b1; e += 64) g |= frontier[nbr[e]];
    return g;
}

// Pull form of one level, for symmetric CSRs (u -> v present iff v -> u): every node
// that some source has not reached ORs the frontier words of its neighbours.  No
// atomics; a node whose unreached sources all appear stops early.  Lane per node of
// degree <= 64, whole wavefront per node of degree <= hub_deg (as in the push); the
// larger ones are k_bfs_hubs' (their next[] word is left at the 0 that apply wrote).
__global__ __launch_bounds__(kBfsBlock) void k_bfs_pull(const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                                                        const uint64_t *__restrict__ frontier,
                                                        const uint64_t *__restrict__ visited, uint64_t *next, uint32_t n,
                                                        uint64_t full, uint32_t hub_deg)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (kBfsBlock / 64);
    for (uint32_t base = (blockIdx.x * (kBfsBlock / 64) + threadIdx.x / 64) * 64; base < n; base += waves * 64) {
        const uint32_t v = base + lane;  // uniform loop bound: every lane takes part in the ballot
        uint64_t want = 0;                // sources that have not reached v
        uint32_t e0 = 0, e1 = 0;
        if (v < n) {
            want = full & ~visited[v];
            if (want) {
                e0 = off[v];
                e1 = off[v + 1];
            }
        }
        uint64_t got = 0;
        if (e1 - e0 <= 64) {
            // 8 arcs per step, loads issued together; the early stop is checked per step
            constexpr uint32_t B = 8;
            for (uint64_t e = e0; e < e1 && (got & want) != want; e += B) {
                uint32_t w[B];
#pragma unroll
                for (uint32_t j = 0; j < B; ++j) w[j] = e + j < e1 ? nbr[e + j] : 0u;
#pragma unroll
                for (uint32_t j = 0; j < B; ++j)
                    if (e + j < e1) got |= frontier[w[j]];
            }
        }
        if (v < n && e1 - e0 <= 64) next[v] = got & want;
        unsigned long long big = __ballot(e1 - e0 > 64 && e1 - e0 <= hub_deg);
        while (big) {
            const int i = __ffsll(big) - 1;
            big &= big - 1;
            const uint32_t b0 = (uint32_t)__shfl((int)e0, i), b1 = (uint32_t)__shfl((int)e1, i);
            const uint64_t wi = (uint64_t)__shfl((long long)want, i);
            uint64_t g = hb_bfs_or_range(nbr, frontier, b0, b1, lane);
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) g |= (uint64_t)__shfl_xor((long long)g, o);
            if ((int)lane == i) next[base + (uint32_t)i] = g & wi;
        }
    }
}

// The lists of nodes of degree > hub_deg, cut into chunks (hub_node[c], arcs
// [hub_beg[c], hub_end[c])), one wavefront per chunk.  Pull: a chunk ORs its frontier
// words and ORs the bits its hub still wants into next[hub] (zero before the level).
// Push: a chunk of a frontier hub pushes the hub's word along its arcs.
template <bool kPull>
__global__ __launch_bounds__(kBfsBlock) void k_bfs_hubs(const uint32_t *__restrict__ nbr,
                                                        const uint32_t *__restrict__ hub_node,
                                                        const uint32_t *__restrict__ hub_beg,
                                                        const uint32_t *__restrict__ hub_end, uint32_t chunks,
                                                        const uint64_t *__restrict__ frontier,
                                                        const uint64_t *__restrict__ visited, uint64_t *next,
                                                        uint64_t full)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (kBfsBlock / 64);
    for (uint32_t c = blockIdx.x * (kBfsBlock / 64) + threadIdx.x / 64; c < chunks; c += waves) {
        const uint32_t v = hub_node[c], b0 = hub_beg[c], b1 = hub_end[c];
        if (kPull) {
            const uint64_t want = full & ~visited[v];
            if (!want) continue;
            uint64_t g = hb_bfs_or_range(nbr, frontier, b0, b1, lane);
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) g |= (uint64_t)__shfl_xor((long long)g, o);
            if (lane == 0 && (g & want)) atomicOr((unsigned long long *)&next[v], (unsigned long long)(g & want));
        } else {
            const uint64_t f = frontier[v];
            if (!f) continue;
            for (uint32_t e = b0 + lane; e < b1; e += 64) hb_bfs_push_arc(visited, next, nbr[e], f);
        }
    }
}

// newly = next & ~visited; visited |= newly; frontier = newly; per-source counts of
// newly (one ballot per bit lane per wavefront) and the arcs of the new frontier's
// nodes: the block's 65 values go to part[65 * block], summed by sum_columns; dist
// (optional, [node][64] u16) records the level of each newly set bit
constexpr uint32_t kBfsCols = 65;

__global__ __launch_bounds__(kBfsBlock) void k_bfs_apply(const uint32_t *__restrict__ off, uint64_t *frontier,
                                                         uint64_t *visited, uint64_t *next, uint16_t *dist, uint32_t n,
                                                         uint32_t level, unsigned long long *part)
{
    __shared__ unsigned long long wc[kBfsBlock / 64][kBfsCols];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    unsigned long long mine = 0;  // count of bit `lane` over this wavefront's nodes
    unsigned long long farcs = 0;  // arcs of this lane's nodes that joined the frontier
    // the loop bound is uniform across the wavefront, so every lane takes part in the ballots
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const uint32_t v = base + threadIdx.x;
        uint64_t newly = 0;
        if (v < n) {
            const uint64_t nx = next[v];
            if (nx) {
                const uint64_t vis = visited[v];
                newly = nx & ~vis;
                next[v] = 0ull;
                if (newly) {
                    visited[v] = vis | newly;
                    farcs += off[v + 1] - off[v];
                    if (dist) {  // the row's 16-byte pieces holding newly set bits, read-modify-write
                        uint4 *row = (uint4 *)(dist + (size_t)v * 64);
#pragma unroll
                        for (int j = 0; j < 8; ++j) {
                            const uint32_t bits = (uint32_t)(newly >> (8 * j)) & 0xffu;
                            if (!bits) continue;
                            uint4 w = row[j];
                            uint32_t q[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                            for (int h = 0; h < 8; ++h) {
                                const uint32_t sh = 16u * (uint32_t)(h & 1);
                                if ((bits >> h) & 1u) q[h >> 1] = (q[h >> 1] & ~(0xffffu << sh)) | (level << sh);
                            }
                            row[j] = make_uint4(q[0], q[1], q[2], q[3]);
                        }
                    }
                }
            }
            frontier[v] = newly;
        }
        if (__ballot(newly != 0ull)) {
            for (int k = 0; k < 64; ++k) {
                const unsigned long long m = __ballot((newly >> k) & 1ull);
                if ((int)lane == k) mine += (unsigned long long)__popcll(m);
            }
        }
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) farcs += (unsigned long long)__shfl_xor((long long)farcs, o);
    wc[wave][lane] = mine;
    if (lane == 0) wc[wave][64] = farcs;
    __syncthreads();
    if (threadIdx.x < kBfsCols) {
        unsigned long long s = 0;
        for (int w = 0; w < kBfsBlock / 64; ++w) s += wc[w][threadIdx.x];
        part[kBfsCols * blockIdx.x + threadIdx.x] = s;
    }
}

// inv_sum[64 b + k] = Σ_{v=0..n-1} inv[dist_b[v][k]] as a sequential f64 sum in node
// order, dist_b = dist + b n 64 the distances of pass b of a group; inv[0] = 0
// (unreached), inv[d] = 1/d for d < entries.  One block per pass: the adds of a source
// are one dependent chain (lane k of wavefront 0 = source k), so the passes of a group
// run side by side, and the whole block streams the distances through LDS in tiles of
// kInvTile nodes -- tile t + 1 is loaded while tile t is summed -- so the summing
// wavefront never waits on HBM.  A 1/d table of at most kInvLds entries sits in LDS.
constexpr uint32_t kInvLds = 4096;
constexpr uint32_t kInvBlock = 256;
constexpr uint32_t kInvTile = 256;                    // nodes per tile: 32 KiB of u16 distances
constexpr uint32_t kInvVec = kInvTile * 64 * 2 / 16;  // uint4 per tile
constexpr uint32_t kInvPer = kInvVec / kInvBlock;     // uint4 per thread per tile

template <bool kLds>
__global__ __launch_bounds__(kInvBlock) void k_bfs_inv_sum(const uint16_t *__restrict__ dist,
                                                           const double *__restrict__ inv, uint32_t entries, uint32_t n,
                                                           const uint32_t *__restrict__ counts, double *out)
{
    __shared__ double tab[kLds ? kInvLds : 1];
    __shared__ uint4 tile[2][kInvVec];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    if (kLds)
        for (uint32_t i = tid; i < entries; i += kInvBlock) tab[i] = inv[i];
    const uint4 *__restrict__ src = (const uint4 *)(dist + (size_t)b * n * 64);  // 8 uint4 per node
    const uint32_t tiles = (n + kInvTile - 1) / kInvTile;
    auto vecs = [&](uint32_t t) -> uint32_t {
        return (uint32_t)(std::min<uint64_t>(n, (uint64_t)(t + 1) * kInvTile) - (uint64_t)t * kInvTile) * 8u;
    };
    uint4 r[kInvPer];
    auto load = [&](uint32_t t) {
        const uint32_t m = vecs(t);
        const uint4 *p = src + (size_t)t * kInvVec;
#pragma unroll
        for (uint32_t j = 0; j < kInvPer; ++j) {
            const uint32_t i = tid + j * kInvBlock;
            if (i < m) r[j] = p[i];
        }
    };
    auto store = [&](uint32_t buf, uint32_t t) {
        const uint32_t m = vecs(t);
#pragma unroll
        for (uint32_t j = 0; j < kInvPer; ++j) {
            const uint32_t i = tid + j * kInvBlock;
            if (i < m) tile[buf][i] = r[j];
        }
    };
    double acc = 0.0;
    if (tiles) {
        load(0);
        store(0, 0);
    }
    __syncthreads();
    for (uint32_t t = 0; t < tiles; ++t) {
        if (t + 1 < tiles) load(t + 1);  // in flight while wavefront 0 sums tile t
        if (tid < 64) {
            const uint16_t *d = (const uint16_t *)tile[t & 1];
            const uint32_t m = vecs(t) / 8u;
            uint32_t j = 0;
            for (; j + 16 <= m; j += 16) {
                double x[16];
#pragma unroll
                for (uint32_t q = 0; q < 16; ++q) {
                    const uint16_t dv = d[(j + q) * 64 + tid];
                    x[q] = kLds ? tab[dv] : inv[dv];
                }
#pragma unroll
                for (uint32_t q = 0; q < 16; ++q) acc = __dadd_rn(acc, x[q]);
            }
            for (; j < m; ++j) {
                const uint16_t dv = d[j * 64 + tid];
                acc = __dadd_rn(acc, kLds ? tab[dv] : inv[dv]);
            }
        }
        if (t + 1 < tiles) store((t + 1) & 1, t + 1);
        __syncthreads();
    }
    if (tid < counts[b]) out[64 * b + tid] = acc;
}

}  // namespace hb

namespace {
int efail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    hb::set_error(buf);
    return code;
}

struct Scoped {
    void *p = nullptr;
    ~Scoped()
    {
        if (p) (void)hipFree(p);
    }
};
}  // namespace

#define HBX_TRY(call)                                                                                    \
    do {                                                                                                 \
        const hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                            \
            return efail(e_ == hipErrorOutOfMemory ? HB_ENOMEM : HB_EHIP, "hb_exact_bfs: %s failed: %s", \
                         #call, hipGetErrorString(e_));                                                  \
    } while (0)

namespace hb {

int exact_bfs(const uint32_t *off, const uint32_t *h_off, const uint32_t *nbr, uint32_t n, bool symmetric,
              uint64_t s_begin,
              uint64_t s_end, uint64_t *reach, uint64_t *total, double *inv_sum, hipStream_t s)
{
    if (s_begin > s_end || s_end > n) return efail(HB_EINVAL, "hb_exact_bfs: sources [%llu, %llu) outside [0, %u)",
                                                   (unsigned long long)s_begin, (unsigned long long)s_end, n);
    if (s_begin == s_end) return HB_OK;
    Scoped bits, cnt, dist, inv, sums, cnts;
    const unsigned grid = (unsigned)std::min<uint64_t>((n + kBfsBlock - 1) / kBfsBlock, 4096);
    // passes of 64 sources; with harmonic sums, the distances of up to `group` passes are
    // kept (u16 per node and source) so their sequential sums run side by side
    const uint64_t passes = (s_end - s_begin + 63) / 64;
    const uint64_t pass_bytes = (uint64_t)n * 64 * 2;
    const uint32_t group = inv_sum ? (uint32_t)std::max<uint64_t>(
                                         1, std::min<uint64_t>({32, passes, (16ull << 30) / pass_bytes}))
                                   : 1;
    HBX_TRY(hipMalloc(&bits.p, (size_t)3 * n * 8));
    HBX_TRY(hipMalloc(&cnt.p, (size_t)kBfsCols * (1 + grid) * 8));  // totals, then kBfsCols per block
    // hubs: nodes of degree > kHubDeg, their lists cut into chunks of kHubChunk arcs
    // (one wavefront each).  A wavefront walks the lists of degree 65..kHubDeg of its
    // 64 nodes one after another, so that range is kept short.
    constexpr uint32_t kHubDeg = 128, kHubChunk = 1024;
    const uint32_t arcs = h_off[n];
    std::vector<uint32_t> h_hub;  // node, beg, end per chunk, as three runs
    {
        std::vector<uint32_t> hn, hb_, he;
        for (uint32_t v = 0; v < n; ++v)
            if (h_off[v + 1] - h_off[v] > kHubDeg)
                for (uint64_t b = h_off[v]; b < h_off[v + 1]; b += kHubChunk) {
                    hn.push_back(v);
                    hb_.push_back((uint32_t)b);
                    he.push_back((uint32_t)std::min<uint64_t>(b + kHubChunk, h_off[v + 1]));
                }
        h_hub = hn;
        h_hub.insert(h_hub.end(), hb_.begin(), hb_.end());
        h_hub.insert(h_hub.end(), he.begin(), he.end());
    }
    const uint32_t chunks = (uint32_t)(h_hub.size() / 3);
    Scoped hub;
    if (chunks) {
        HBX_TRY(hipMalloc(&hub.p, h_hub.size() * 4));
        HBX_TRY(hipMemcpy(hub.p, h_hub.data(), h_hub.size() * 4, hipMemcpyHostToDevice));
    }
    const uint32_t *hub_node = (const uint32_t *)hub.p, *hub_beg = hub_node + chunks, *hub_end = hub_beg + chunks;
    const unsigned hub_grid = (unsigned)std::min<uint64_t>((chunks + 3) / 4, 4096);
    if (inv_sum) {
        HBX_TRY(hipMalloc(&dist.p, (size_t)group * pass_bytes));
        HBX_TRY(hipMalloc(&sums.p, (size_t)group * 64 * 8));
        HBX_TRY(hipMalloc(&cnts.p, (size_t)group * 4));
    }
    uint64_t *fr = (uint64_t *)bits.p, *vis = fr + n, *nx = fr + 2 * (size_t)n;
    unsigned long long *d_cnt = (unsigned long long *)cnt.p;
    std::vector<unsigned long long> hc(kBfsCols);
    std::vector<double> h_inv;
    std::vector<uint32_t> h_counts(group);
    for (uint64_t g0 = 0; g0 < passes; g0 += group) {
        const uint32_t gn = (uint32_t)std::min<uint64_t>(group, passes - g0);
        uint32_t levels = 0;
        for (uint32_t j = 0; j < gn; ++j) {
            const uint64_t s0 = s_begin + (g0 + j) * 64;
            const uint32_t count = (uint32_t)std::min<uint64_t>(64, s_end - s0);
            uint16_t *dj = dist.p ? (uint16_t *)dist.p + (size_t)j * n * 64 : nullptr;
            h_counts[j] = count;
            uint64_t r[64] = {}, t[64] = {};
            const uint64_t full = count == 64 ? ~0ull : ((1ull << count) - 1);
            // arcs leaving the frontier: the sources' own lists first
            uint64_t frontier_arcs = h_off[s0 + count] - h_off[s0];
            k_bfs_init<<<grid, kBfsBlock, 0, s>>>(fr, vis, nx, n, (uint32_t)s0, count);
            if (dj) k_bfs_zero<uint16_t><<<4096, kBfsBlock, 0, s>>>(dj, (uint64_t)n * 64);
            for (uint32_t level = 1;; ++level) {
                if (dj && level > 65535u)
                    return efail(HB_ERANGE, "hb_exact_bfs: eccentricity above 65535 (harmonic sums need u16 distances)");
                // direction (symmetric graphs only): pull when the frontier's arcs exceed
                // 1/8 of all arcs -- a pull scans every unreached node's list once without
                // atomics, a push walks the frontier's lists with an atomic OR per arc
                if (level == 1) {
                    const unsigned sg = (unsigned)std::max<uint64_t>(
                        1, std::min<uint64_t>((frontier_arcs + kBfsBlock - 1) / kBfsBlock, 4096));
                    k_bfs_push_sources<<<sg, kBfsBlock, 0, s>>>(off, nbr, nx, (uint32_t)s0, count);
                } else if (symmetric && frontier_arcs * 8 > arcs) {
                    k_bfs_pull<<<grid, kBfsBlock, 0, s>>>(off, nbr, fr, vis, nx, n, full, kHubDeg);
                    if (chunks)
                        k_bfs_hubs<true><<<hub_grid, kBfsBlock, 0, s>>>(nbr, hub_node, hub_beg, hub_end, chunks, fr,
                                                                        vis, nx, full);
                } else {
                    k_bfs_push<<<grid, kBfsBlock, 0, s>>>(off, nbr, fr, vis, nx, n, kHubDeg);
                    if (chunks)
                        k_bfs_hubs<false><<<hub_grid, kBfsBlock, 0, s>>>(nbr, hub_node, hub_beg, hub_end, chunks, fr,
                                                                         vis, nx, full);
                }
                k_bfs_apply<<<grid, kBfsBlock, 0, s>>>(off, fr, vis, nx, dj, n, level, d_cnt + kBfsCols);
                HBX_TRY(sum_columns(d_cnt + kBfsCols, grid, kBfsCols, d_cnt, s));
                HBX_TRY(hipMemcpyAsync(hc.data(), d_cnt, kBfsCols * 8, hipMemcpyDeviceToHost, s));
                HBX_TRY(hipStreamSynchronize(s));
                bool any = false;
                for (int k = 0; k < 64; ++k) any = any || hc[k] != 0;
                frontier_arcs = hc[64];
                if (!any) break;
                levels = std::max(levels, level);
                for (int k = 0; k < 64; ++k) {
                    r[k] += hc[k];
                    t[k] += (uint64_t)level * hc[k];
                }
            }
            for (uint32_t k = 0; k < count; ++k) {
                if (reach) reach[s0 - s_begin + k] = r[k];
                if (total) total[s0 - s_begin + k] = t[k];
            }
        }
        if (inv_sum) {
            // 1/d as correctly rounded host divisions, d = 1..levels; entry 0 is an unreached node
            if (h_inv.size() < (size_t)levels + 1) {
                const size_t from = h_inv.empty() ? 1 : h_inv.size();
                h_inv.resize((size_t)levels + 1);
                h_inv[0] = 0.0;
                for (size_t d = from; d <= levels; ++d) h_inv[d] = 1.0 / (double)d;
                if (inv.p) (void)hipFree(inv.p), inv.p = nullptr;
                HBX_TRY(hipMalloc(&inv.p, h_inv.size() * 8));
                HBX_TRY(hipMemcpyAsync(inv.p, h_inv.data(), h_inv.size() * 8, hipMemcpyHostToDevice, s));
            }
            HBX_TRY(hipMemcpyAsync(cnts.p, h_counts.data(), (size_t)gn * 4, hipMemcpyHostToDevice, s));
            const uint32_t entries = (uint32_t)h_inv.size();
            if (entries <= kInvLds)
                k_bfs_inv_sum<true><<<gn, kInvBlock, 0, s>>>((const uint16_t *)dist.p, (const double *)inv.p, entries, n,
                                                      (const uint32_t *)cnts.p, (double *)sums.p);
            else
                k_bfs_inv_sum<false><<<gn, kInvBlock, 0, s>>>((const uint16_t *)dist.p, (const double *)inv.p, entries, n,
                                                       (const uint32_t *)cnts.p, (double *)sums.p);
            HBX_TRY(hipGetLastError());
            std::vector<double> hs((size_t)gn * 64);
            HBX_TRY(hipMemcpyAsync(hs.data(), sums.p, hs.size() * 8, hipMemcpyDeviceToHost, s));
            HBX_TRY(hipStreamSynchronize(s));
            for (uint32_t j = 0; j < gn; ++j)
                for (uint32_t k = 0; k < h_counts[j]; ++k)
                    inv_sum[(g0 + j) * 64 + k] = hs[(size_t)j * 64 + k];
        }
    }
    return HB_OK;
}

}  // namespace hb
