// sparkx_hb.hpp -- header-only C++ mirror of the reference's HyperBallInner
// (src/centralities/hyper_ball.rs:81-108, aliases HyperBall4..16 :46-58) over the
// C ABI of libsparkx_hb.so (sparkx_hb.h).  Same constructor and method names,
// argument meaning and error behaviour: fallible calls throw HyperBallError (the
// reference returns Err), out-of-range node getters fail an assertion-like check
// (the reference's assert!, :262,:268), and centralities are cached per
// normalisation as in the reference (:290-424).
#pragma once

#include <cstdint>
#include <cstdio>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "sparkx_hb.h"

namespace sparkx {

struct HyperBallError : std::runtime_error {
    int code;
    HyperBallError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc)
{
    if (rc < 0) throw HyperBallError(rc, hb_last_error());
}

// Borrowed CSR view (GraphMemoryMap's index and edge arrays; 4- or 8-byte entries).
struct CsrView {
    const void *offsets = nullptr;
    int offset_bytes = 8;
    uint64_t nodes = 0;
    const void *neighbours = nullptr;
    int neighbour_bytes = 8;
    uint64_t arcs = 0;
    uint64_t size() const { return nodes; }   // GraphMemoryMap::size
    uint64_t width() const { return arcs; }   // GraphMemoryMap::width
};

template <int P>
class HyperBall {
    static_assert(P >= 4 && P <= 16, "precision 4..16");

public:
    // HyperBallInner::new (hyper_ball.rs:124-133)
    static HyperBall create(const CsrView &g, int device = 0) { return HyperBall(g, 0, 0u, device); }
    // HyperBallInner::new_with_conf (:144-157): max_depth and threads do not change
    // results in the reference (max_t is dead code, threads only split its CPU loop)
    static HyperBall new_with_conf(const CsrView &g, std::optional<size_t> /*max_depth*/,
                                   size_t /*threads*/, int device = 0)
    {
        return HyperBall(g, 0, 0u, device);
    }
    // HyperBallInner::get_or_compute (:165-201): reopens the registers
    // (hyperballregisters_<P>_<id>.mmap, this library's own file: the reference's
    // counters file holds hyperloglog-rs structs whose layout is not available) and the
    // accumulators (hyperballdistances_<id>.mmap, hyperballinvdistances_<id>.mmap) from
    // cache_dir, then searches the centrality caches; without them it runs and saves
    // them.  The object keeps the cache: compute_* write their files there.  A
    // reopened run reports 0 iterations, as the reference's does.
    static HyperBall get_or_compute(const CsrView &g, const std::string &cache_dir, const std::string &gid,
                                    int device = 0)
    {
        const size_t n = (size_t)g.nodes;
        std::vector<uint8_t> regs;
        std::vector<double> dist, inv;
        if (read_file(file(cache_dir, "hyperballregisters", gid, P), regs, n << P) &&
            read_file(file(cache_dir, "hyperballdistances", gid, -1), dist, n) &&
            read_file(file(cache_dir, "hyperballinvdistances", gid, -1), inv, n)) {
            HyperBall h(g, device);
            check(hb_restore(h.ctx_, P, regs.data(), dist.data(), inv.data()));
            h.stats_ = std::make_unique<hb_stats>();  // zeros: 0 iterations
            h.stats_->p = P;
            h.cache_dir_ = cache_dir;
            h.gid_ = gid;
            // every cached centrality that exists (the reference ignores the others)
            const int kinds[5][2] = {{HB_CLOSENESS, 0}, {HB_HARMONIC, 0}, {HB_CLOSENESS, 1}, {HB_HARMONIC, 1},
                                     {HB_LIN, 0}};
            for (const auto &k : kinds) {
                std::vector<double> &slot = k[0] == HB_LIN ? h.lin_ : k[0] == HB_CLOSENESS ? h.closeness_[k[1]]
                                                                                            : h.harmonic_[k[1]];
                try {
                    (void)h.search(k[0], k[1] != 0, slot);
                } catch (const HyperBallError &) {
                }
            }
            return h;
        }
        HyperBall h(g, 0, 0u, device);
        h.cache_dir_ = cache_dir;
        h.gid_ = gid;
        write_file(file(cache_dir, "hyperballregisters", gid, P), h.counters());
        write_file(file(cache_dir, "hyperballdistances", gid, -1), h.geodesic_distances());
        write_file(file(cache_dir, "hyperballinvdistances", gid, -1), h.harmonic_geodesic_distances());
        return h;
    }

    // extension: truncation and scheduling flags (HB_FLAG_*)
    HyperBall(const CsrView &g, int max_iters, uint32_t flags, int device = 0) : HyperBall(g, device)
    {
        stats_ = std::make_unique<hb_stats>();
        check(hb_run(ctx_, P, max_iters, flags, stats_.get()));
    }
    HyperBall(const HyperBall &) = delete;
    HyperBall &operator=(const HyperBall &) = delete;
    HyperBall(HyperBall &&o) noexcept { *this = std::move(o); }
    HyperBall &operator=(HyperBall &&o) noexcept
    {
        std::swap(ctx_, o.ctx_);
        std::swap(g_, o.g_);
        std::swap(stats_, o.stats_);
        std::swap(closeness_, o.closeness_);
        std::swap(harmonic_, o.harmonic_);
        std::swap(lin_, o.lin_);
        std::swap(cache_dir_, o.cache_dir_);
        std::swap(gid_, o.gid_);
        return *this;
    }
    ~HyperBall()
    {
        if (ctx_) hb_destroy(ctx_);
    }

    // iterations executed, including the final one without increase (:557-560)
    size_t get_iters() const { return (size_t)stats_->iterations; }
    const hb_stats &stats() const { return *stats_; }

    // :272-280
    std::vector<double> geodesic_distances() const
    {
        std::vector<double> d(g_.nodes);
        check(hb_read_accumulators(ctx_, d.data(), nullptr));
        return d;
    }
    std::vector<double> harmonic_geodesic_distances() const
    {
        std::vector<double> d(g_.nodes);
        check(hb_read_accumulators(ctx_, nullptr, d.data()));
        return d;
    }
    // :260-270 (assert!(node_id < self.g.size()))
    double get_geodesic_distance(size_t node_id) const
    {
        if (node_id >= g_.nodes) throw std::out_of_range("node_id >= |V|");
        return geodesic_distances()[node_id];
    }
    double get_harmonic_geodesic_distance(size_t node_id) const
    {
        if (node_id >= g_.nodes) throw std::out_of_range("node_id >= |V|");
        return harmonic_geodesic_distances()[node_id];
    }

    // :434-477, :487-529, :532-555 (written to the cache when the object has one)
    const std::vector<double> &compute_closeness_centrality(bool normalize)
    {
        return saved(HB_CLOSENESS, normalize, closeness_[normalize] = centrality(HB_CLOSENESS, normalize));
    }
    const std::vector<double> &compute_harmonic_centrality(bool normalize)
    {
        return saved(HB_HARMONIC, normalize, harmonic_[normalize] = centrality(HB_HARMONIC, normalize));
    }
    const std::vector<double> &compute_lins_centrality()
    {
        return saved(HB_LIN, false, lin_ = centrality(HB_LIN, false));
    }

    // :302-317, :359-374, :405-415: the cached file into its slot, else Err
    const std::vector<double> &search_cache_closeness_centrality(bool normalization)
    {
        return search(HB_CLOSENESS, normalization, closeness_[normalization]);
    }
    const std::vector<double> &search_cache_harmonic_centrality(bool normalization)
    {
        return search(HB_HARMONIC, normalization, harmonic_[normalization]);
    }
    const std::vector<double> &search_cache_lins_centrality() { return search(HB_LIN, false, lin_); }

    // :290-299, :347-356, :397-402: Err("no such cached centrality")
    const std::vector<double> &get_closeness_centrality(bool normalization) const
    {
        return cached(closeness_[normalization]);
    }
    const std::vector<double> &get_hamonic_centrality(bool normalization) const
    {
        return cached(harmonic_[normalization]);
    }
    const std::vector<double> &get_lins_centrality() const { return cached(lin_); }

    // :327-337, :384-394, :418-424.  The closeness variant tests the *harmonic* slot
    // (:332), reproduced: empty -> recompute, else the cached closeness (throws if none).
    const std::vector<double> &get_or_compute_closeness_centrality(bool normalization)
    {
        return harmonic_[normalization].empty() ? compute_closeness_centrality(normalization)
                                                : get_closeness_centrality(normalization);
    }
    const std::vector<double> &get_or_compute_harmonic_centrality(bool normalization)
    {
        return harmonic_[normalization].empty() && g_.nodes ? compute_harmonic_centrality(normalization)
                                                            : harmonic_[normalization];
    }
    const std::vector<double> &get_or_compute_lins_centrality()
    {
        return lin_.empty() && g_.nodes ? compute_lins_centrality() : lin_;
    }

    // the counters (u8 registers, node-major) the reference leaves in its file
    std::vector<uint8_t> counters(int which = HB_REGS_REFERENCE) const
    {
        int run_p = 0;  // the engine's completed run must be this mirror's precision
        check(hb_result_info(ctx_, &run_p, nullptr, nullptr));
        if (run_p != P) throw HyperBallError(HB_ESTATE, "the context holds a run of another precision");
        std::vector<uint8_t> r(g_.nodes << P);
        check(hb_read_registers(ctx_, which, r.data()));
        return r;
    }
    std::vector<float> estimates(int which = HB_REGS_REFERENCE) const
    {
        std::vector<float> e(g_.nodes);
        check(hb_read_estimates(ctx_, which, e.data()));
        return e;
    }

private:
    // a context with the graph staged, no run yet
    HyperBall(const CsrView &g, int device) : g_(g)
    {
        check(hb_create(&ctx_, device));
        const int rc = hb_stage_csr(ctx_, g.offsets, g.offset_bytes, g.nodes, g.neighbours, g.neighbour_bytes, g.arcs);
        if (rc < 0) {
            const HyperBallError e(rc, hb_last_error());
            hb_destroy(ctx_);
            ctx_ = nullptr;
            throw e;
        }
    }

    // <suffix>_<id>.mmap, or <suffix>_<seq>_<id>.mmap (graph/cache/utils.rs:318-341)
    static std::string file(const std::string &dir, const char *suffix, const std::string &gid, int seq)
    {
        return dir + "/" + suffix + (seq >= 0 ? "_" + std::to_string(seq) : std::string()) + "_" + gid + ".mmap";
    }
    // closeness / harmonic: seq 1 without normalisation, 2 with; Lin 0 (hyper_ball.rs:203-246)
    std::string centrality_file(int kind, bool normalize) const
    {
        const char *suffix = kind == HB_CLOSENESS ? "hyperballcloseness" : kind == HB_HARMONIC ? "hyperballharmonic"
                                                                                              : "hyperballlin";
        return file(cache_dir_, suffix, gid_, kind == HB_LIN ? 0 : (normalize ? 2 : 1));
    }
    template <class T>
    static bool read_file(const std::string &path, std::vector<T> &out, size_t count)
    {
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) return false;
        out.resize(count);
        const size_t got = count ? std::fread(out.data(), sizeof(T), count, f) : 0;
        const bool at_end = std::fgetc(f) == EOF;
        std::fclose(f);
        return got == count && at_end;
    }
    template <class T>
    static void write_file(const std::string &path, const std::vector<T> &v)
    {
        FILE *f = std::fopen(path.c_str(), "wb");
        if (!f) throw HyperBallError(HB_EINVAL, "cannot write " + path);
        const size_t put = v.empty() ? 0 : std::fwrite(v.data(), sizeof(T), v.size(), f);
        std::fclose(f);
        if (put != v.size()) throw HyperBallError(HB_EINVAL, "short write to " + path);
    }
    const std::vector<double> &saved(int kind, bool normalize, const std::vector<double> &v) const
    {
        if (!cache_dir_.empty()) write_file(centrality_file(kind, normalize), v);
        return v;
    }
    const std::vector<double> &search(int kind, bool normalize, std::vector<double> &slot)
    {
        std::vector<double> v;
        if (cache_dir_.empty() || !read_file(centrality_file(kind, normalize), v, (size_t)g_.nodes))
            throw HyperBallError(HB_ESTATE, "error no such cached centrality");
        return slot = std::move(v);
    }

    std::vector<double> centrality(int kind, bool normalize) const
    {
        std::vector<double> c(g_.nodes);
        check(hb_centrality(ctx_, kind, normalize ? 1 : 0, c.data()));
        return c;
    }
    static const std::vector<double> &cached(const std::vector<double> &v)
    {
        if (v.empty()) throw HyperBallError(HB_ESTATE, "error no such cached centrality");
        return v;
    }

    hb_ctx *ctx_ = nullptr;
    CsrView g_{};
    std::unique_ptr<hb_stats> stats_;
    std::vector<double> closeness_[2], harmonic_[2], lin_;
    std::string cache_dir_, gid_;  // set when opened through get_or_compute
};

using HyperBall4 = HyperBall<4>;
using HyperBall5 = HyperBall<5>;
using HyperBall6 = HyperBall<6>;
using HyperBall7 = HyperBall<7>;
using HyperBall8 = HyperBall<8>;
using HyperBall9 = HyperBall<9>;
using HyperBall10 = HyperBall<10>;
using HyperBall11 = HyperBall<11>;
using HyperBall12 = HyperBall<12>;
using HyperBall13 = HyperBall<13>;
using HyperBall14 = HyperBall<14>;
using HyperBall15 = HyperBall<15>;
using HyperBall16 = HyperBall<16>;

}  // namespace sparkx
