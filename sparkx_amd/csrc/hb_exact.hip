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
// reference searches them.
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
// serialise on one lane.
__device__ __forceinline__ void hb_bfs_push_arc(const uint64_t *__restrict__ visited, uint64_t *next, uint32_t v,
                                                uint64_t f)
{
    const uint64_t add = f & ~visited[v];
    if (add) atomicOr((unsigned long long *)&next[v], (unsigned long long)add);
}

__global__ __launch_bounds__(kBfsBlock) void k_bfs_push(const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                                                        const uint64_t *__restrict__ frontier,
                                                        const uint64_t *__restrict__ visited, uint64_t *next, uint32_t n)
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
        unsigned long long big = __ballot(e1 - e0 > 64);
        while (big) {
            const int i = __ffsll(big) - 1;
            big &= big - 1;
            const uint32_t b0 = (uint32_t)__shfl((int)e0, i), b1 = (uint32_t)__shfl((int)e1, i);
            const uint64_t fi = (uint64_t)__shfl((long long)f, i);
            for (uint32_t e = b0 + lane; e < b1; e += 64) hb_bfs_push_arc(visited, next, nbr[e], fi);
        }
    }
}

// newly = next & ~visited; visited |= newly; frontier = newly; per-source counts of
// newly (one ballot per bit lane per wavefront; the block's 64 counts go to
// part[64 * block], summed by sum_columns); dist (optional, [node][64] u16)
// records the level of each newly set bit
__global__ __launch_bounds__(kBfsBlock) void k_bfs_apply(uint64_t *frontier, uint64_t *visited, uint64_t *next,
                                                         uint16_t *dist, uint32_t n, uint32_t level,
                                                         unsigned long long *part)
{
    __shared__ unsigned long long wc[kBfsBlock / 64][64];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    unsigned long long mine = 0;  // count of bit `lane` over this wavefront's nodes
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
                    if (dist) {
                        uint64_t b = newly;
                        while (b) {
                            const int k = __ffsll((unsigned long long)b) - 1;
                            dist[(size_t)v * 64 + k] = (uint16_t)level;
                            b &= b - 1;
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
    wc[wave][lane] = mine;
    __syncthreads();
    if (threadIdx.x < 64) {
        unsigned long long s = 0;
        for (int w = 0; w < kBfsBlock / 64; ++w) s += wc[w][threadIdx.x];
        part[64 * blockIdx.x + threadIdx.x] = s;
    }
}

// inv_sum[k] = Σ_{v=0..n-1} inv[dist[v][k]] as a sequential f64 sum in node order;
// one wavefront, lane k = source k; inv[0] = 0 (unreached), inv[d] = 1/d
__global__ __launch_bounds__(64) void k_bfs_inv_sum(const uint16_t *__restrict__ dist, const double *__restrict__ inv,
                                                    uint32_t n, uint32_t count, double *out)
{
    const uint32_t k = threadIdx.x;
    constexpr uint32_t U = 16;
    double acc = 0.0;
    uint32_t v = 0;
    for (; v + U <= n; v += U) {
        double t[U];
#pragma unroll
        for (uint32_t j = 0; j < U; ++j) t[j] = inv[dist[(size_t)(v + j) * 64 + k]];
#pragma unroll
        for (uint32_t j = 0; j < U; ++j) acc = __dadd_rn(acc, t[j]);
    }
    for (; v < n; ++v) acc = __dadd_rn(acc, inv[dist[(size_t)v * 64 + k]]);
    if (k < count) out[k] = acc;
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

int exact_bfs(const uint32_t *off, const uint32_t *nbr, uint32_t n, uint64_t s_begin, uint64_t s_end,
              uint64_t *reach, uint64_t *total, double *inv_sum, hipStream_t s)
{
    if (s_begin > s_end || s_end > n) return efail(HB_EINVAL, "hb_exact_bfs: sources [%llu, %llu) outside [0, %u)",
                                                   (unsigned long long)s_begin, (unsigned long long)s_end, n);
    if (s_begin == s_end) return HB_OK;
    Scoped bits, cnt, dist, inv, sums;
    const unsigned grid = (unsigned)std::min<uint64_t>((n + kBfsBlock - 1) / kBfsBlock, 4096);
    HBX_TRY(hipMalloc(&bits.p, (size_t)3 * n * 8));
    HBX_TRY(hipMalloc(&cnt.p, (size_t)(64 + 64 * grid) * 8));  // 64 totals, then 64 per block
    if (inv_sum) {
        HBX_TRY(hipMalloc(&dist.p, (size_t)n * 64 * 2));
        HBX_TRY(hipMalloc(&sums.p, 64 * 8));
    }
    uint64_t *fr = (uint64_t *)bits.p, *vis = fr + n, *nx = fr + 2 * (size_t)n;
    unsigned long long *d_cnt = (unsigned long long *)cnt.p;
    std::vector<unsigned long long> hc(64);
    std::vector<double> h_inv;
    for (uint64_t s0 = s_begin; s0 < s_end; s0 += 64) {
        const uint32_t count = (uint32_t)std::min<uint64_t>(64, s_end - s0);
        uint64_t r[64] = {}, t[64] = {};
        uint32_t levels = 0;
        k_bfs_init<<<grid, kBfsBlock, 0, s>>>(fr, vis, nx, n, (uint32_t)s0, count);
        if (dist.p) k_bfs_zero<uint16_t><<<4096, kBfsBlock, 0, s>>>((uint16_t *)dist.p, (uint64_t)n * 64);
        for (uint32_t level = 1;; ++level) {
            if (dist.p && level > 65535u)
                return efail(HB_ERANGE, "hb_exact_bfs: eccentricity above 65535 (harmonic sums need u16 distances)");
            k_bfs_push<<<grid, kBfsBlock, 0, s>>>(off, nbr, fr, vis, nx, n);
            k_bfs_apply<<<grid, kBfsBlock, 0, s>>>(fr, vis, nx, (uint16_t *)dist.p, n, level, d_cnt + 64);
            HBX_TRY(sum_columns(d_cnt + 64, grid, 64, d_cnt, s));
            HBX_TRY(hipMemcpyAsync(hc.data(), d_cnt, 64 * 8, hipMemcpyDeviceToHost, s));
            HBX_TRY(hipStreamSynchronize(s));
            bool any = false;
            for (int k = 0; k < 64; ++k) any = any || hc[k] != 0;
            if (!any) break;
            levels = level;
            for (int k = 0; k < 64; ++k) {
                r[k] += hc[k];
                t[k] += (uint64_t)level * hc[k];
            }
        }
        for (uint32_t k = 0; k < count; ++k) {
            if (reach) reach[s0 - s_begin + k] = r[k];
            if (total) total[s0 - s_begin + k] = t[k];
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
            k_bfs_inv_sum<<<1, 64, 0, s>>>((const uint16_t *)dist.p, (const double *)inv.p, n, count, (double *)sums.p);
            HBX_TRY(hipMemcpyAsync(inv_sum + (s0 - s_begin), sums.p, (size_t)count * 8, hipMemcpyDeviceToHost, s));
            HBX_TRY(hipStreamSynchronize(s));
        }
    }
    return HB_OK;
}

}  // namespace hb
