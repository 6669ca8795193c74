// hb_graphgen.cpp -- seeded synthetic graph generators (see include/sparkx_hb_graphgen.h).
//
// All randomness is counter-based (splitmix64 of seed and sample index), so the
// edge multiset depends only on the arguments; the CSR is then canonical
// (symmetric, sorted, deduplicated, loop-free) and independent of the thread count.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <functional>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sparkx_hb_graphgen.h"

struct hb_graph {
    uint64_t n = 0;
    std::vector<uint32_t> off, nbr;
};

namespace {

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {  // splitmix64 stream
    uint64_t x;
    explicit Rng(uint64_t seed, uint64_t stream) : x(mix64(seed ^ mix64(stream + 0x632BE59BD9B4E019ULL))) {}
    uint64_t next()
    {
        x += 0x9E3779B97F4A7C15ULL;
        return mix64(x);
    }
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

int clamp_threads(int t)
{
    if (t > 0) return t;
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 64u));
}

// f(begin, end) over [0, n) in `threads` contiguous ranges
void parallel_ranges(uint64_t n, int threads, const std::function<void(uint64_t, uint64_t)> &f)
{
    if (threads <= 1 || n < 4096) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    const uint64_t step = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const uint64_t b = std::min(n, t * step), e = std::min(n, b + step);
        if (b < e) th.emplace_back(f, b, e);
    }
    for (auto &x : th) x.join();
}

// dynamic schedule over [0, n) in blocks of `grain`
void parallel_dynamic(uint64_t n, uint64_t grain, int threads, const std::function<void(uint64_t, uint64_t)> &f)
{
    std::atomic<uint64_t> next{0};
    auto work = [&]() {
        for (;;) {
            const uint64_t b = next.fetch_add(grain);
            if (b >= n) break;
            f(b, std::min(n, b + grain));
        }
    };
    if (threads <= 1) {
        work();
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < threads; ++t) th.emplace_back(work);
    for (auto &x : th) x.join();
}

// CSR from an undirected edge list: both directions, self-loops dropped,
// each adjacency sorted and deduplicated.
int build_csr(hb_graph *g, uint64_t n, const std::vector<uint32_t> &src, const std::vector<uint32_t> &dst,
              int threads)
{
    const uint64_t S = src.size();
    std::vector<uint32_t> deg(n + 1, 0);
    parallel_ranges(S, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            if (src[i] == dst[i]) continue;
            __atomic_fetch_add(&deg[src[i]], 1u, __ATOMIC_RELAXED);
            __atomic_fetch_add(&deg[dst[i]], 1u, __ATOMIC_RELAXED);
        }
    });
    std::vector<uint64_t> start(n + 1, 0);
    for (uint64_t u = 0; u < n; ++u) start[u + 1] = start[u] + deg[u];
    std::vector<uint64_t> pos(start.begin(), start.end() - 1);
    std::vector<uint32_t> tmp(start[n]);
    parallel_ranges(S, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; ++i) {
            const uint32_t s = src[i], d = dst[i];
            if (s == d) continue;
            tmp[__atomic_fetch_add(&pos[s], 1ull, __ATOMIC_RELAXED)] = d;
            tmp[__atomic_fetch_add(&pos[d], 1ull, __ATOMIC_RELAXED)] = s;
        }
    });
    std::vector<uint64_t>().swap(pos);
    parallel_dynamic(n, 1024, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t u = b; u < e; ++u) {
            uint32_t *lo = tmp.data() + start[u], *hi = tmp.data() + start[u + 1];
            std::sort(lo, hi);
            deg[u] = (uint32_t)(std::unique(lo, hi) - lo);
        }
    });
    g->n = n;
    g->off.assign(n + 1, 0);
    uint64_t total = 0;
    for (uint64_t u = 0; u < n; ++u) {
        g->off[u] = (uint32_t)total;
        total += deg[u];
        if (total >= (1ull << 32)) return -5;
    }
    g->off[n] = (uint32_t)total;
    g->nbr.resize(total);
    parallel_dynamic(n, 4096, threads, [&](uint64_t b, uint64_t e) {
        for (uint64_t u = b; u < e; ++u)
            std::memcpy(g->nbr.data() + g->off[u], tmp.data() + start[u], (size_t)deg[u] * 4);
    });
    return 0;
}

template <class F>
int guarded(hb_graph **out, F &&f)
{
    if (!out) return -1;
    *out = nullptr;
    hb_graph *g = nullptr;
    try {
        g = new hb_graph();
        const int rc = f(g);
        if (rc) {
            delete g;
            return rc;
        }
    } catch (const std::bad_alloc &) {
        delete g;
        return -2;
    }
    *out = g;
    return 0;
}

}  // namespace

extern "C" {

int hb_graph_er(hb_graph **out, uint64_t n, uint64_t samples, uint64_t seed, int threads)
{
    if (n == 0 || n >= (1ull << 32) - 1) return -1;
    threads = clamp_threads(threads);
    return guarded(out, [&](hb_graph *g) {
        std::vector<uint32_t> src(samples), dst(samples);
        parallel_ranges(samples, threads, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i) {
                Rng r(seed, i);
                src[i] = (uint32_t)r.below(n);
                dst[i] = (uint32_t)r.below(n);
            }
        });
        return build_csr(g, n, src, dst, threads);
    });
}

int hb_graph_rmat(hb_graph **out, int scale, uint64_t samples, double a, double b, double c,
                  uint64_t seed, int threads)
{
    if (scale < 1 || scale > 31 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return -1;
    threads = clamp_threads(threads);
    const uint64_t n = 1ull << scale;
    const uint64_t ta = (uint64_t)(a * 4294967296.0), tab = (uint64_t)((a + b) * 4294967296.0),
                   tabc = (uint64_t)((a + b + c) * 4294967296.0);
    return guarded(out, [&](hb_graph *g) {
        std::vector<uint32_t> src(samples), dst(samples);
        parallel_ranges(samples, threads, [&](uint64_t bb, uint64_t e) {
            for (uint64_t i = bb; i < e; ++i) {
                Rng r(seed, i);
                uint32_t s = 0, d = 0;
                uint64_t bits = 0;
                for (int l = 0; l < scale; ++l) {
                    if ((l & 1) == 0) bits = r.next();
                    const uint64_t x = (l & 1) ? (bits >> 32) : (bits & 0xffffffffull);
                    const uint32_t qs = x >= tab, qd = (x >= ta && x < tab) || x >= tabc;
                    s = (s << 1) | qs;
                    d = (d << 1) | qd;
                }
                src[i] = s;
                dst[i] = d;
            }
        });
        return build_csr(g, n, src, dst, threads);
    });
}

int hb_graph_kmer(hb_graph **out, uint64_t genome_len, int copies, double sub_rate, int k,
                  uint64_t seed, int threads)
{
    if (k < 2 || k > 31 || copies < 1 || genome_len < (uint64_t)k || sub_rate < 0 || sub_rate > 1) return -1;
    threads = clamp_threads(threads);
    return guarded(out, [&](hb_graph *g) {
        const uint64_t L = genome_len, P = L - (uint64_t)k + 1;  // k-mers per copy
        std::vector<uint8_t> genome(L);
        parallel_ranges((L + 31) / 32, threads, [&](uint64_t b, uint64_t e) {
            for (uint64_t w = b; w < e; ++w) {
                Rng r(seed, w);
                const uint64_t x = r.next();
                for (uint64_t j = 0; j < 32 && 32 * w + j < L; ++j) genome[32 * w + j] = (x >> (2 * j)) & 3;
            }
        });
        const uint64_t kmask = (k == 32) ? ~0ull : ((1ull << (2 * k)) - 1);
        const uint64_t thr = (uint64_t)(sub_rate * 9007199254740992.0);  // 2^53
        std::vector<uint64_t> codes((uint64_t)copies * P);
        for (int cpy = 0; cpy < copies; ++cpy) {
            std::vector<uint8_t> seq(genome);
            parallel_ranges(L, threads, [&](uint64_t b, uint64_t e) {
                for (uint64_t i = b; i < e; ++i) {
                    Rng r(seed ^ 0x5DEECE66DULL, (uint64_t)cpy * L + i);
                    const uint64_t x = r.next();
                    if ((x >> 11) < thr) seq[i] = (uint8_t)((seq[i] + 1 + (r.next() % 3)) & 3);
                }
            });
            parallel_ranges(P, threads, [&](uint64_t b, uint64_t e) {
                uint64_t fw = 0, rc = 0;
                for (uint64_t i = b; i < b + (uint64_t)k - 1; ++i) {
                    fw = ((fw << 2) | seq[i]) & kmask;
                    rc = (rc >> 2) | ((uint64_t)(3 - seq[i]) << (2 * (k - 1)));
                }
                for (uint64_t i = b; i < e; ++i) {
                    const uint8_t base = seq[i + k - 1];
                    fw = ((fw << 2) | base) & kmask;
                    rc = (rc >> 2) | ((uint64_t)(3 - base) << (2 * (k - 1)));
                    codes[(uint64_t)cpy * P + i] = std::min(fw, rc);
                }
            });
        }
        // distinct canonical k-mers in ascending order: chunk sorts + pairwise merges
        std::vector<uint64_t> uniq(codes);
        {
            const int T = threads;
            const uint64_t N = uniq.size(), step = (N + T - 1) / T;
            std::vector<uint64_t> bounds;
            for (int t = 0; t <= T; ++t) bounds.push_back(std::min(N, t * step));
            parallel_dynamic((uint64_t)T, 1, T, [&](uint64_t b, uint64_t e) {
                for (uint64_t t = b; t < e; ++t) std::sort(uniq.begin() + bounds[t], uniq.begin() + bounds[t + 1]);
            });
            for (uint64_t width = 1; width < (uint64_t)T; width *= 2) {
                std::vector<std::thread> th;
                for (uint64_t t = 0; t + width < (uint64_t)T; t += 2 * width) {
                    const uint64_t lo = bounds[t], mid = bounds[t + width],
                                   hi = bounds[std::min<uint64_t>(T, t + 2 * width)];
                    th.emplace_back([&uniq, lo, mid, hi]() {
                        std::inplace_merge(uniq.begin() + lo, uniq.begin() + mid, uniq.begin() + hi);
                    });
                }
                for (auto &x : th) x.join();
            }
            uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
        }
        const uint64_t n = uniq.size();
        if (n >= (1ull << 32) - 1) return -5;
        std::vector<uint32_t> id(codes.size());
        parallel_ranges(codes.size(), threads, [&](uint64_t b, uint64_t e) {
            for (uint64_t i = b; i < e; ++i)
                id[i] = (uint32_t)(std::lower_bound(uniq.begin(), uniq.end(), codes[i]) - uniq.begin());
        });
        std::vector<uint64_t>().swap(codes);
        const uint64_t E = (uint64_t)copies * (P - 1);
        std::vector<uint32_t> src(E), dst(E);
        for (int cpy = 0; cpy < copies; ++cpy)
            for (uint64_t i = 0; i + 1 < P; ++i) {
                src[(uint64_t)cpy * (P - 1) + i] = id[(uint64_t)cpy * P + i];
                dst[(uint64_t)cpy * (P - 1) + i] = id[(uint64_t)cpy * P + i + 1];
            }
        return build_csr(g, n, src, dst, threads);
    });
}

uint64_t hb_graph_nodes(const hb_graph *g) { return g ? g->n : 0; }
uint64_t hb_graph_arcs(const hb_graph *g) { return g ? g->nbr.size() : 0; }
const uint32_t *hb_graph_offsets(const hb_graph *g) { return g ? g->off.data() : nullptr; }
const uint32_t *hb_graph_neighbours(const hb_graph *g) { return g ? g->nbr.data() : nullptr; }
void hb_graph_free(hb_graph *g) { delete g; }

}  // extern "C"
