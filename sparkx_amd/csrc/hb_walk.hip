// hb_walk.hip -- walk-count inputs of the learned surrogate.
//
// For every node u: S_0(u) = deg(u), S_k(u) = Σ over u's arcs u -> v of S_{k-1}(v)
// (the number of walks of k + 1 arcs leaving u), k = 1..passes, and M_1(u) = the
// largest degree among u's heads.  Sums are u64 and saturate at 2^64 - 1: a saturating
// sum of non-negative terms is min(2^64 - 1, exact sum) whatever the order of the
// adds, and a maximum is order-free too, so every partition of a list over lanes,
// wavefronts and chunks gives the same bits.
//
// One pass per k reads S_{k-1} at every arc's head (pass 1 the u32 degrees, half the
// bytes of a u64 row).  Layout of a pass, as in the exact searches' pull levels: a
// wavefront takes 64 consecutive nodes; a node of degree <= 64 is summed by its own
// lane, one of degree <= kWalkHub by the whole wavefront, and a larger list is cut
// into kWalkHub-arc chunks (one wavefront each) whose partial sums are combined per
// hub in a second launch.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include <hip/hip_runtime.h>

#include "hb_kernels.h"

namespace hb {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ uint64_t sat_add(uint64_t a, uint64_t b)
{
    const uint64_t r = a + b;
    return r < a ? ~0ull : r;
}

// this lane's share (e = b0 + lane, b0 + lane + 64, ... < b1) of a list's sum and max
template <typename In>
__device__ __forceinline__ void walk_range(const uint32_t *__restrict__ nbr, const In *__restrict__ in,
                                           uint64_t b0, uint64_t b1, uint32_t lane, uint64_t &s, uint64_t &m)
{
    uint64_t e = b0 + lane;
    for (; e + 3 * 64 < b1; e += 4 * 64) {
        const uint32_t w0 = nbr[e], w1 = nbr[e + 64], w2 = nbr[e + 128], w3 = nbr[e + 192];
        const uint64_t x0 = in[w0], x1 = in[w1], x2 = in[w2], x3 = in[w3];
        s = sat_add(sat_add(s, x0), sat_add(x1, sat_add(x2, x3)));
        m = std::max(std::max(m, x0), std::max(std::max(x1, x2), x3));
    }
    for (; e < b1; e += 64) {
        const uint64_t x = in[nbr[e]];
        s = sat_add(s, x);
        m = std::max(m, x);
    }
}

__device__ __forceinline__ void wave_reduce(uint64_t &s, uint64_t &m)
{
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        s = sat_add(s, (uint64_t)__shfl_xor((long long)s, o));
        m = std::max(m, (uint64_t)__shfl_xor((long long)m, o));
    }
}

// out[u] = Σ in[v], mx[u] = max in[v] (mx nullable) over u's arcs, for every node of
// degree <= hub_deg; the larger ones are k_walk_chunks' and k_walk_hubs'
template <typename In>
__global__ __launch_bounds__(kBlock) void k_walk(const uint32_t *__restrict__ off, const uint32_t *__restrict__ nbr,
                                                 uint32_t n, uint32_t hub_deg, const In *__restrict__ in,
                                                 uint64_t *out, uint64_t *mx)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (kBlock / 64);
    for (uint32_t base = (blockIdx.x * (kBlock / 64) + threadIdx.x / 64) * 64; base < n; base += waves * 64) {
        const uint32_t v = base + lane;  // uniform loop bound: every lane takes part in the ballot
        uint32_t e0 = 0, e1 = 0;
        if (v < n) {
            e0 = off[v];
            e1 = off[v + 1];
        }
        if (v < n && e1 - e0 <= 64) {
            uint64_t s = 0, m = 0;
            uint32_t e = e0;
            for (; e + 4 <= e1; e += 4) {
                const uint32_t w0 = nbr[e], w1 = nbr[e + 1], w2 = nbr[e + 2], w3 = nbr[e + 3];
                const uint64_t x0 = in[w0], x1 = in[w1], x2 = in[w2], x3 = in[w3];
                s = sat_add(sat_add(s, x0), sat_add(x1, sat_add(x2, x3)));
                m = std::max(std::max(m, x0), std::max(std::max(x1, x2), x3));
            }
            for (; e < e1; ++e) {
                const uint64_t x = in[nbr[e]];
                s = sat_add(s, x);
                m = std::max(m, x);
            }
            out[v] = s;
            if (mx) mx[v] = m;
        }
        unsigned long long big = __ballot(e1 - e0 > 64 && e1 - e0 <= hub_deg);
        while (big) {
            const int i = __ffsll(big) - 1;
            big &= big - 1;
            const uint32_t b0 = (uint32_t)__shfl((int)e0, i), b1 = (uint32_t)__shfl((int)e1, i);
            uint64_t s = 0, m = 0;
            walk_range(nbr, in, b0, b1, lane, s, m);
            wave_reduce(s, m);
            if ((int)lane == i) {
                out[base + (uint32_t)i] = s;
                if (mx) mx[base + (uint32_t)i] = m;
            }
        }
    }
}

// chunk c covers arcs [beg[c], end[c]) of one hub: its sum and max into psum / pmax
template <typename In>
__global__ __launch_bounds__(kBlock) void k_walk_chunks(const uint32_t *__restrict__ nbr,
                                                        const uint32_t *__restrict__ beg,
                                                        const uint32_t *__restrict__ end, uint32_t chunks,
                                                        const In *__restrict__ in, uint64_t *psum, uint64_t *pmax)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t waves = gridDim.x * (kBlock / 64);
    for (uint32_t c = blockIdx.x * (kBlock / 64) + threadIdx.x / 64; c < chunks; c += waves) {
        uint64_t s = 0, m = 0;
        walk_range(nbr, in, beg[c], end[c], lane, s, m);
        wave_reduce(s, m);
        if (lane == 0) {
            psum[c] = s;
            pmax[c] = m;
        }
    }
}

// hub h owns chunks [first[h], first[h + 1]): combined into out / mx of node node[h]
__global__ __launch_bounds__(kBlock) void k_walk_hubs(const uint32_t *__restrict__ node,
                                                      const uint32_t *__restrict__ first, uint32_t hubs,
                                                      const uint64_t *__restrict__ psum,
                                                      const uint64_t *__restrict__ pmax, uint64_t *out, uint64_t *mx)
{
    const uint32_t h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= hubs) return;
    uint64_t s = 0, m = 0;
    for (uint32_t c = first[h]; c < first[h + 1]; ++c) {
        s = sat_add(s, psum[c]);
        m = std::max(m, pmax[c]);
    }
    out[node[h]] = s;
    if (mx) mx[node[h]] = m;
}

__global__ __launch_bounds__(kBlock) void k_degrees(const uint32_t *__restrict__ off, uint32_t n, uint32_t *deg)
{
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += gridDim.x * blockDim.x)
        deg[u] = off[u + 1] - off[u];
}

// the inputs of node u, L = f64 log1p of the exact (or saturated) integer, each
// value rounded to f32 once:
//   2 passes: L(deg), L(S_1), L(S_2), L(M_1), L(S_2) - L(S_1), L(S_1) - L(deg)
//   3 passes: L(deg), L(S_1), L(S_2), L(S_3), L(M_1), L(S_3) - L(S_2)
__global__ __launch_bounds__(kBlock) void k_walk_inputs(const uint32_t *__restrict__ deg,
                                                        const uint64_t *__restrict__ s, uint32_t n, uint32_t x_nodes,
                                                        int passes, float *x)
{
    for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < x_nodes; u += gridDim.x * blockDim.x) {
        const double ld = log1p((double)deg[u]);
        const double l1 = log1p((double)s[u]);
        const double l2 = log1p((double)s[(uint64_t)n + u]);
        const double lm = log1p((double)s[3 * (uint64_t)n + u]);
        float *r = x + (uint64_t)u * 6;
        r[0] = (float)ld;
        r[1] = (float)l1;
        r[2] = (float)l2;
        if (passes == 3) {
            const double l3 = log1p((double)s[2 * (uint64_t)n + u]);
            r[3] = (float)l3;
            r[4] = (float)lm;
            r[5] = (float)(l3 - l2);
        } else {
            r[3] = (float)lm;
            r[4] = (float)(l2 - l1);
            r[5] = (float)(l1 - ld);
        }
    }
}

unsigned grid_for(uint64_t items)
{
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + kBlock - 1) / kBlock, 16384));
}

}  // namespace

hipError_t walk_inputs(const uint32_t *off, const uint32_t *nbr, uint32_t n, const WalkHubs &hubs, int passes,
                       uint64_t *s, uint32_t *deg, uint64_t *part, float *x, uint32_t x_nodes, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    if (passes != 2 && passes != 3) return hipErrorInvalidValue;
    uint64_t *m1 = s + 3 * (uint64_t)n;
    const unsigned grid = grid_for((uint64_t)n);  // 256 nodes per block: a wavefront per 64
    const unsigned cgrid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((hubs.chunks + 3) / 4, 4096));
    const unsigned hgrid = (hubs.hubs + kBlock - 1) / kBlock;
    k_degrees<<<grid, kBlock, 0, st>>>(off, n, deg);
    k_walk<uint32_t><<<grid, kBlock, 0, st>>>(off, nbr, n, kWalkHub, deg, s, m1);
    if (hubs.chunks) {
        k_walk_chunks<uint32_t><<<cgrid, kBlock, 0, st>>>(nbr, hubs.beg, hubs.end, hubs.chunks, deg, part,
                                                           part + hubs.chunks);
        k_walk_hubs<<<hgrid, kBlock, 0, st>>>(hubs.node, hubs.first, hubs.hubs, part, part + hubs.chunks, s, m1);
    }
    for (int k = 2; k <= passes; ++k) {
        const uint64_t *in = s + (uint64_t)(k - 2) * n;
        uint64_t *out = s + (uint64_t)(k - 1) * n;
        k_walk<uint64_t><<<grid, kBlock, 0, st>>>(off, nbr, n, kWalkHub, in, out, nullptr);
        if (hubs.chunks) {
            k_walk_chunks<uint64_t><<<cgrid, kBlock, 0, st>>>(nbr, hubs.beg, hubs.end, hubs.chunks, in, part,
                                                               part + hubs.chunks);
            k_walk_hubs<<<hgrid, kBlock, 0, st>>>(hubs.node, hubs.first, hubs.hubs, part, part + hubs.chunks, out,
                                                  nullptr);
        }
    }
    if (x_nodes) k_walk_inputs<<<grid_for(x_nodes), kBlock, 0, st>>>(deg, s, n, x_nodes, passes, x);
    return hipGetLastError();
}

}  // namespace hb
