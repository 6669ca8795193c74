// hb_batch.hip -- many independent HyperBall runs (graphs of different sizes and
// precisions p = 4..8) advanced together: one launch per iteration of a
// persistent work-queue kernel over (graph, degree-bin) tiles, every graph with
// its own iteration state, termination rule and buffer parity.  The iteration
// (begin / union / finalize) is capturable in one hipGraph and replayed until
// every graph has stopped.  Each graph's registers, estimates and accumulators
// are bit-identical to its own single-graph run (same per-node union_node).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/sparkx_hb.h"
#include "hb_device.h"
#include "hb_kernels.h"
#include "hb_union.h"


namespace hb {

static constexpr int kBatchBlock = 256;
static constexpr int kBatchPMin = 4, kBatchPMax = 8;
// LC tables of p = 4..8 packed back to back: offsets of each p's table
static constexpr int kLcOff[kBatchPMax + 1] = {0, 0, 0, 0, 0, 17, 17 + 33, 17 + 33 + 65, 17 + 33 + 65 + 129};
static constexpr int kLcTotal = 17 + 33 + 65 + 129 + 257;

struct HbBatchGraph {
    uint64_t off_base, nbr_base, reg_base, node_base, bits_base, rec_base;
    uint32_t n, p, n_linked, words, rec_cap, overflow;
    float am, five_m;
    double lc_thr;
    HbIterState st;
};

struct HbBatchTile {
    uint32_t graph, first, count, team;
};

struct HbBatchDev {
    const uint32_t *off, *nbr;
    uint8_t *regs0, *regs1;
    float *est;
    double *dist, *inv;
    uint32_t *bits0, *bits1;
    HbBatchGraph *g;
    uint32_t ngraphs;
    const HbBatchTile *tiles;
    uint32_t ntiles;
    const uint32_t *lists;  // graph-local node ids, tiles index into it
    HbIterRec *rec;
    uint32_t *queue;        // work-queue head of the current iteration
    const float *lc_all;    // LC tables for p = 4..8 (kLcOff)
    const HbBiasTable *bias_all; // bias windows of p = 4..8
    int frontier;
};

__global__ __launch_bounds__(kBatchBlock) void k_batch_begin(HbBatchDev b)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *b.queue = 0u;
    for (uint32_t gi = blockIdx.x; gi < b.ngraphs; gi += gridDim.x) {
        const HbBatchGraph &g = b.g[gi];
        if (g.st.done) continue;
        uint32_t *bw = ((g.st.iter & 1u) ? b.bits1 : b.bits0) + g.bits_base;
        for (uint32_t i = threadIdx.x; i < g.words; i += blockDim.x) bw[i] = 0u;
    }
}

template <int P, int T>
__device__ __forceinline__ void batch_tile(const HbBatchDev &b, const HbBatchGraph &g, const HbBatchTile &tile,
                                           const HbTables &tb, HbCounts &c)
{
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1);
    const uint32_t it = g.st.iter;
    HbView v;
    v.off = b.off + g.off_base;
    v.nbr = b.nbr + g.nbr_base;
    v.orig = v.nbr;
    v.node_id = nullptr;
    v.in = (g.st.parity ? b.regs1 : b.regs0) + g.reg_base;
    v.out = (g.st.parity ? b.regs0 : b.regs1) + g.reg_base;
    v.est = b.est + g.node_base;
    v.dist = b.dist + g.node_base;
    v.inv = b.inv + g.node_base;
    v.bitsR = ((it & 1u) ? b.bits0 : b.bits1) + g.bits_base;
    v.bitsW = ((it & 1u) ? b.bits1 : b.bits0) + g.bits_base;
    v.t = g.st.t;
    v.inv_t = g.st.inv_t;
    v.use_frontier = b.frontier && it > 1 && !g.st.skip_scan;
    v.first = false;
    v.gather_bits = nullptr;
    v.scr = nullptr;
    const uint32_t *list = b.lists + tile.first;
    for (uint32_t i = lane / T; i < tile.count; i += HB_WAVE / T) union_node<P, T>(v, tb, list[i], lane % T, c);
}

template <int P>
__device__ __forceinline__ void batch_tile_p(const HbBatchDev &b, const HbBatchGraph &g, const HbBatchTile &tile,
                                             const HbTables &tb, HbCounts &c)
{
    constexpr int G = (1 << P) / 16;
    switch (tile.team) {
    case 1: if constexpr (G <= 1) batch_tile<P, 1>(b, g, tile, tb, c); break;
    case 2: if constexpr (G <= 2) batch_tile<P, 2>(b, g, tile, tb, c); break;
    case 4: if constexpr (G <= 4) batch_tile<P, 4>(b, g, tile, tb, c); break;
    case 8: if constexpr (G <= 8) batch_tile<P, 8>(b, g, tile, tb, c); break;
    case 16: batch_tile<P, 16>(b, g, tile, tb, c); break;
    case 32: batch_tile<P, 32>(b, g, tile, tb, c); break;
    default: batch_tile<P, 64>(b, g, tile, tb, c); break;
    }
}

__global__ __launch_bounds__(kBatchBlock) void k_batch_union(HbBatchDev b)
{
    __shared__ float s_lc[kLcTotal];
    __shared__ HbBiasTable s_bias[kBatchPMax - kBatchPMin + 1];
    for (int i = threadIdx.x; i < kLcTotal; i += blockDim.x) s_lc[i] = b.lc_all[i];
    for (int i = threadIdx.x; i < (int)(sizeof s_bias / 8); i += blockDim.x)
        ((uint64_t *)s_bias)[i] = ((const uint64_t *)b.bias_all)[i];
    __syncthreads();
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1);
    for (;;) {
        uint32_t ti = 0;
        if (lane == 0) ti = atomicAdd(b.queue, 1u);
        ti = __shfl(ti, 0);
        if (ti >= b.ntiles) break;
        const HbBatchTile tile = b.tiles[ti];
        const HbBatchGraph &g = b.g[tile.graph];
        if (g.st.done) continue;
        const int p = (int)g.p;
        const HbTables tb = {s_lc + kLcOff[p], &s_bias[p - kBatchPMin], g.am, g.five_m, g.lc_thr};
        HbCounts c;
        switch (p) {
        case 4: batch_tile_p<4>(b, g, tile, tb, c); break;
        case 5: batch_tile_p<5>(b, g, tile, tb, c); break;
        case 6: batch_tile_p<6>(b, g, tile, tb, c); break;
        case 7: batch_tile_p<7>(b, g, tile, tb, c); break;
        default: batch_tile_p<8>(b, g, tile, tb, c); break;
        }
#pragma unroll
        for (int s = 1; s < HB_WAVE; s <<= 1) {
            c.inc += __shfl_xor(c.inc, s);
            c.arcs += __shfl_xor(c.arcs, s);
            c.nodes += __shfl_xor(c.nodes, s);
            c.chg += __shfl_xor(c.chg, s);
        }
        if (lane == 0) {
            HbIterRec *r = b.rec + g.rec_base + g.st.iter;
            if (c.inc) atomicAdd(&r->inc[0], c.inc);
            if (c.arcs) atomicAdd(&r->arcs[0], c.arcs);
            if (c.nodes) atomicAdd(&r->nodes[0], c.nodes);
            if (c.chg) atomicAdd(&r->chg[0], c.chg);
        }
    }
}

// end of an iteration for every running graph (hyper_ball.rs:833-849)
__global__ void k_batch_finalize(HbBatchDev b)
{
    const uint32_t gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= b.ngraphs) return;
    HbBatchGraph &g = b.g[gi];
    HbIterState &st = g.st;
    if (st.done) return;
    const uint32_t it = st.iter;
    const HbIterRec &r = b.rec[g.rec_base + it];
    const unsigned long long inc = r.inc[0], chg = r.chg[0];
    if (inc == 0 || (st.max_iters && it >= st.max_iters) || it + 2 >= g.rec_cap) {
        st.done = 1;
        st.executed = it;
        st.converged = inc == 0;
        g.overflow = inc != 0 && !(st.max_iters && it >= st.max_iters);
        return;
    }
    st.iter = it + 1;
    st.parity ^= 1u;
    st.t = st.t + 1.0;
    st.inv_t = __ddiv_rn(1.0, st.t);
    st.skip_scan = 2 * chg > (unsigned long long)g.n_linked;
}

}  // namespace hb

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

int bfail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    hb::set_error(buf);
    return code;
}

#define HB_BTRY(call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return bfail(e_ == hipErrorOutOfMemory ? HB_ENOMEM : HB_EHIP, "%s failed: %s", #call, \
                         hipGetErrorString(e_));                                               \
    } while (0)

const double kLcThr[9] = {0, 0, 0, 0, 10, 20, 40, 80, 220};

float alpha_f32(uint32_t m)
{
    if (m == 16) return 0.673f;
    if (m == 32) return 0.697f;
    if (m == 64) return 0.709f;
    return 0.7213f / (1.0f + 1.079f / (float)m);
}

template <class T>
T *dmalloc(size_t count, hipError_t *err)
{
    T *p = nullptr;
    *err = hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T));
    return *err == hipSuccess ? p : nullptr;
}

}  // namespace

struct hb_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<hb::HbBatchGraph> hg;       // host copy (bases, constants)
    std::vector<std::vector<uint32_t>> h_off;
    std::vector<int> max_iters;
    uint64_t total_nodes = 0, total_arcs = 0, total_regs = 0, total_words = 0, total_rec = 0;
    uint32_t *off = nullptr, *nbr = nullptr, *bits0 = nullptr, *bits1 = nullptr, *queue = nullptr;
    uint32_t *lists = nullptr;
    uint8_t *regs0 = nullptr, *regs1 = nullptr;
    float *est = nullptr, *lc_all = nullptr, *est_tmp = nullptr;
    double *dist = nullptr, *inv = nullptr;
    HbBiasTable *bias_all = nullptr;
    float *lc_p[9] = {};                   // per-p LC tables for the single-graph launchers
    hb::HbBatchGraph *g = nullptr;
    hb::HbBatchTile *tiles = nullptr;
    uint32_t ntiles = 0;
    HbIterRec *rec = nullptr;
    uint32_t small_max = 8, mid_max = 64, tile_arcs = 16384;
    bool sched_ok = false, ran = false;
    std::vector<hb::HbBatchGraph> h_final;
    std::vector<HbIterRec> h_rec;

    ~hb_batch()
    {
        void *ptrs[] = {off, nbr, bits0, bits1, queue, lists, regs0, regs1, est, lc_all, est_tmp,
                        dist, inv, bias_all, g, tiles, rec};
        for (void *p : ptrs)
            if (p) (void)hipFree(p);
        for (float *p : lc_p)
            if (p) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static HbDev graph_dev(hb_batch *b, int gi)
{
    const hb::HbBatchGraph &g = b->hg[gi];
    HbDev d{};
    d.off = b->off + g.off_base;
    d.nbr = b->nbr + g.nbr_base;
    d.regs0 = b->regs0 + g.reg_base;
    d.regs1 = b->regs1 + g.reg_base;
    d.est = b->est + g.node_base;
    d.dist = b->dist + g.node_base;
    d.inv = b->inv + g.node_base;
    d.lc = b->lc_p[g.p];
    d.bias = b->bias_all + (g.p - hb::kBatchPMin);
    d.n = g.n;
    d.n_linked = g.n_linked;
    d.p = g.p;
    d.am = g.am;
    d.five_m = g.five_m;
    d.lc_thr = g.lc_thr;
    return d;
}

static int build_tiles(hb_batch *b)
{
    if (b->sched_ok) return HB_OK;
    std::vector<uint32_t> lists;
    struct T {
        hb::HbBatchTile t;
        uint64_t work;
    };
    std::vector<T> tiles;
    for (size_t gi = 0; gi < b->hg.size(); ++gi) {
        const auto &off = b->h_off[gi];
        const uint32_t p = b->hg[gi].p, G = (1u << p) / 16u;
        const uint32_t teams[3] = {G, std::min(64u, 4 * G), 64u};
        std::vector<uint32_t> bins[3];
        for (uint32_t u = 0; u < b->hg[gi].n; ++u) {
            const uint32_t d = off[u + 1] - off[u];
            if (d == 0) continue;
            bins[d <= b->small_max ? 0 : (d <= b->mid_max ? 1 : 2)].push_back(u);
        }
        for (int k = 0; k < 3; ++k) {
            size_t i = 0;
            const uint32_t per_wave = 64u / teams[k];
            while (i < bins[k].size()) {
                T tt{};
                tt.t.graph = (uint32_t)gi;
                tt.t.first = (uint32_t)lists.size();
                tt.t.team = teams[k];
                uint64_t arcs = 0;
                uint32_t cnt = 0;
                while (i < bins[k].size() && (arcs < b->tile_arcs || cnt % per_wave) && cnt < 4096) {
                    const uint32_t u = bins[k][i++];
                    lists.push_back(u);
                    arcs += off[u + 1] - off[u] + 8;
                    ++cnt;
                }
                tt.t.count = cnt;
                tt.work = arcs << p;
                tiles.push_back(tt);
            }
        }
    }
    // longest tiles first: the dynamic queue then balances the tail
    std::stable_sort(tiles.begin(), tiles.end(), [](const T &a, const T &c) { return a.work > c.work; });
    std::vector<hb::HbBatchTile> ht(tiles.size());
    for (size_t i = 0; i < tiles.size(); ++i) ht[i] = tiles[i].t;
    if (b->lists) (void)hipFree(b->lists);
    if (b->tiles) (void)hipFree(b->tiles);
    b->lists = nullptr;
    b->tiles = nullptr;
    hipError_t e;
    b->lists = dmalloc<uint32_t>(lists.size(), &e);
    HB_BTRY(e);
    b->tiles = dmalloc<hb::HbBatchTile>(ht.size(), &e);
    HB_BTRY(e);
    if (!lists.empty()) HB_BTRY(hipMemcpy(b->lists, lists.data(), lists.size() * 4, hipMemcpyHostToDevice));
    if (!ht.empty()) HB_BTRY(hipMemcpy(b->tiles, ht.data(), ht.size() * sizeof(hb::HbBatchTile), hipMemcpyHostToDevice));
    b->ntiles = (uint32_t)ht.size();
    b->sched_ok = true;
    return HB_OK;
}

extern "C" {

int hb_batch_create(hb_batch **out, int device, const hb_graph_desc *graphs, int count)
{
    if (!out || !graphs || count <= 0) return bfail(HB_EINVAL, "hb_batch_create: bad arguments");
    *out = nullptr;
    int ndev = 0;
    HB_BTRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return bfail(HB_EINVAL, "hb_batch_create: device %d out of range", device);
    HB_BTRY(hipSetDevice(device));
    auto b = std::make_unique<hb_batch>();
    b->device = device;
    HB_BTRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    b->hg.resize(count);
    b->h_off.resize(count);
    b->max_iters.resize(count);
    uint64_t nodes = 0, arcs = 0, regs = 0, words = 0, recs = 0;
    for (int i = 0; i < count; ++i) {
        const hb_graph_desc &gd = graphs[i];
        if (gd.p < hb::kBatchPMin || gd.p > hb::kBatchPMax)
            return bfail(HB_ERANGE, "hb_batch_create: graph %d has p=%d; batches take p = 4..8", i, gd.p);
        if (!gd.offsets || (gd.arcs && !gd.nbrs) || (gd.off_bytes != 4 && gd.off_bytes != 8) ||
            (gd.nbr_bytes != 4 && gd.nbr_bytes != 8) || gd.max_iters < 0 || gd.n == 0 || gd.n >= (1ull << 31) ||
            gd.arcs >= (1ull << 32))
            return bfail(HB_EINVAL, "hb_batch_create: graph %d has a bad descriptor", i);
        auto at = [&](uint64_t k) -> uint64_t {
            return gd.off_bytes == 8 ? ((const uint64_t *)gd.offsets)[k] : ((const uint32_t *)gd.offsets)[k];
        };
        if (at(0) != 0 || at(gd.n) != gd.arcs) return bfail(HB_EINVAL, "hb_batch_create: graph %d offsets do not span its arcs", i);
        auto &ho = b->h_off[i];
        ho.resize(gd.n + 1);
        for (uint64_t k = 0; k <= gd.n; ++k) {
            ho[k] = (uint32_t)at(k);
            if (k && ho[k] < ho[k - 1]) return bfail(HB_EINVAL, "hb_batch_create: graph %d offsets decrease", i);
        }
        for (uint64_t k = 0; k < gd.arcs; ++k) {
            const uint64_t v = gd.nbr_bytes == 8 ? ((const uint64_t *)gd.nbrs)[k] : ((const uint32_t *)gd.nbrs)[k];
            if (v >= gd.n) return bfail(HB_EINVAL, "hb_batch_create: graph %d neighbour %llu >= n", i, (unsigned long long)v);
        }
        hb::HbBatchGraph &g = b->hg[i];
        memset(&g, 0, sizeof g);
        g.off_base = nodes + i;  // n+1 offsets per graph
        g.nbr_base = arcs;
        g.reg_base = regs;
        g.node_base = nodes;
        g.bits_base = words;
        g.rec_base = recs;
        g.n = (uint32_t)gd.n;
        g.p = (uint32_t)gd.p;
        g.words = (uint32_t)((gd.n + 31) / 32);
        g.rec_cap = gd.max_iters > 0 ? (uint32_t)gd.max_iters + 2 : (uint32_t)std::min<uint64_t>(gd.n + 3, 1u << 16);
        const uint32_t m = 1u << gd.p;
        g.am = alpha_f32(m) * ((float)m * (float)m);
        g.five_m = 5.0f * (float)m;
        g.lc_thr = kLcThr[gd.p];
        for (uint64_t k = 0; k < gd.n; ++k) g.n_linked += ho[k + 1] != ho[k];
        b->max_iters[i] = gd.max_iters;
        nodes += gd.n;
        arcs += gd.arcs;
        regs += gd.n << gd.p;
        words += g.words;
        recs += g.rec_cap;
    }
    b->total_nodes = nodes;
    b->total_arcs = arcs;
    b->total_regs = regs;
    b->total_words = words;
    b->total_rec = recs;
    hipError_t e;
    b->off = dmalloc<uint32_t>(nodes + count, &e); HB_BTRY(e);
    b->nbr = dmalloc<uint32_t>(arcs, &e); HB_BTRY(e);
    b->regs0 = dmalloc<uint8_t>(regs, &e); HB_BTRY(e);
    b->regs1 = dmalloc<uint8_t>(regs, &e); HB_BTRY(e);
    b->est = dmalloc<float>(nodes, &e); HB_BTRY(e);
    b->est_tmp = dmalloc<float>(nodes, &e); HB_BTRY(e);
    b->dist = dmalloc<double>(nodes, &e); HB_BTRY(e);
    b->inv = dmalloc<double>(nodes, &e); HB_BTRY(e);
    b->bits0 = dmalloc<uint32_t>(words, &e); HB_BTRY(e);
    b->bits1 = dmalloc<uint32_t>(words, &e); HB_BTRY(e);
    b->queue = dmalloc<uint32_t>(1, &e); HB_BTRY(e);
    b->g = dmalloc<hb::HbBatchGraph>(count, &e); HB_BTRY(e);
    b->rec = dmalloc<HbIterRec>(recs, &e); HB_BTRY(e);
    b->lc_all = dmalloc<float>(hb::kLcTotal, &e); HB_BTRY(e);
    b->bias_all = dmalloc<HbBiasTable>(hb::kBatchPMax - hb::kBatchPMin + 1, &e); HB_BTRY(e);
    std::vector<float> lc(hb::kLcTotal, 0.0f);
    for (int p = hb::kBatchPMin; p <= hb::kBatchPMax; ++p) {
        const uint32_t m = 1u << p;
        std::vector<float> one(m + 1, 0.0f);
        for (uint32_t z = 1; z <= m; ++z) one[z] = (float)((double)m * std::log((double)m / (double)z));
        std::copy(one.begin(), one.end(), lc.begin() + hb::kLcOff[p]);
        b->lc_p[p] = dmalloc<float>(m + 1, &e);
        HB_BTRY(e);
        HB_BTRY(hipMemcpy(b->lc_p[p], one.data(), (m + 1) * 4, hipMemcpyHostToDevice));
        HB_BTRY(hipMemcpy(b->bias_all + (p - hb::kBatchPMin), hb::bias_table(p), sizeof(HbBiasTable),
                          hipMemcpyHostToDevice));
    }
    HB_BTRY(hipMemcpy(b->lc_all, lc.data(), lc.size() * 4, hipMemcpyHostToDevice));
    // CSR upload (per graph, converted to u32)
    std::vector<uint32_t> tmp;
    for (int i = 0; i < count; ++i) {
        const hb_graph_desc &gd = graphs[i];
        const hb::HbBatchGraph &g = b->hg[i];
        HB_BTRY(hipMemcpy(b->off + g.off_base, b->h_off[i].data(), (gd.n + 1) * 4, hipMemcpyHostToDevice));
        tmp.resize(gd.arcs);
        for (uint64_t k = 0; k < gd.arcs; ++k)
            tmp[k] = (uint32_t)(gd.nbr_bytes == 8 ? ((const uint64_t *)gd.nbrs)[k] : ((const uint32_t *)gd.nbrs)[k]);
        if (gd.arcs) HB_BTRY(hipMemcpy(b->nbr + g.nbr_base, tmp.data(), gd.arcs * 4, hipMemcpyHostToDevice));
    }
    *out = b.release();
    hb::set_error("");
    return HB_OK;
}

void hb_batch_destroy(hb_batch *b)
{
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    delete b;
}

int hb_batch_set_schedule(hb_batch *b, uint32_t small_max_deg, uint32_t mid_max_deg, uint32_t tile_arcs)
{
    if (!b) return bfail(HB_EINVAL, "hb_batch_set_schedule: NULL batch");
    if (small_max_deg) b->small_max = small_max_deg;
    if (mid_max_deg) b->mid_max = std::max(mid_max_deg, b->small_max);
    if (tile_arcs) b->tile_arcs = tile_arcs;
    b->sched_ok = false;
    return HB_OK;
}

// The batched engine is switched off: its first GPU run (64 graphs of 4K-262K nodes,
// p = 4..8) stalled without finishing and the cause has not been found by reading the
// code, so it stays off until one is shown.  hb_batch_create and the readers still
// validate their arguments; hb_batch_run (and the autotuner that calls it) return
// HB_ESTATE.
static constexpr bool kBatchEngineEnabled = false;

int hb_batch_run(hb_batch *b, uint32_t flags, hb_batch_stats *stats)
{
    if (!b) return bfail(HB_EINVAL, "hb_batch_run: NULL batch");
    if (!kBatchEngineEnabled)
        return bfail(HB_ESTATE, "hb_batch_run: the batched engine is switched off in this build "
                                "(an unexplained stall of its first GPU run is under investigation)");
    HB_BTRY(hipSetDevice(b->device));
    {
        const int rc = build_tiles(b);
        if (rc != HB_OK) return rc;
    }
    b->ran = false;
    const int ng = (int)b->hg.size();
    std::vector<hb::HbBatchGraph> init = b->hg;
    for (int i = 0; i < ng; ++i) {
        HbIterState &st = init[i].st;
        st = HbIterState{};
        st.iter = 1;
        st.max_iters = (uint32_t)b->max_iters[i];
        st.t = 1.0;
        st.inv_t = 1.0;
        init[i].overflow = 0;
    }
    hb::HbBatchDev d{};
    d.off = b->off;
    d.nbr = b->nbr;
    d.regs0 = b->regs0;
    d.regs1 = b->regs1;
    d.est = b->est;
    d.dist = b->dist;
    d.inv = b->inv;
    d.bits0 = b->bits0;
    d.bits1 = b->bits1;
    d.g = b->g;
    d.ngraphs = (uint32_t)ng;
    d.tiles = b->tiles;
    d.ntiles = b->ntiles;
    d.lists = b->lists;
    d.rec = b->rec;
    d.queue = b->queue;
    d.lc_all = b->lc_all;
    d.bias_all = b->bias_all;
    d.frontier = (flags & HB_FLAG_DENSE) ? 0 : 1;

    hipEvent_t e0, e1, e2;
    HB_BTRY(hipEventCreate(&e0));
    HB_BTRY(hipEventCreate(&e1));
    HB_BTRY(hipEventCreate(&e2));
    int rc = HB_OK;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    uint32_t launches = 0;
    const unsigned grid = (unsigned)std::min<uint32_t>(std::max<uint32_t>(b->ntiles, 1), 4096);
    auto launch_iter = [&]() -> hipError_t {
        hb::k_batch_begin<<<(unsigned)std::min(ng, 1024), hb::kBatchBlock, 0, b->stream>>>(d);
        hb::k_batch_union<<<grid, hb::kBatchBlock, 0, b->stream>>>(d);
        hb::k_batch_finalize<<<(ng + 63) / 64, 64, 0, b->stream>>>(d);
        return hipGetLastError();
    };
    do {
        if (hipMemcpyAsync(b->g, init.data(), ng * sizeof(hb::HbBatchGraph), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
            hipMemsetAsync(b->rec, 0, b->total_rec * sizeof(HbIterRec), b->stream) != hipSuccess ||
            hb::launch_fill(b->bits0, b->total_words, 0u, b->stream) != hipSuccess ||
            hb::launch_fill(b->bits1, b->total_words, 0u, b->stream) != hipSuccess ||
            hipEventRecord(e0, b->stream) != hipSuccess) {
            rc = bfail(HB_EHIP, "hb_batch_run: set-up failed");
            break;
        }
        for (int i = 0; i < ng && rc == HB_OK; ++i)
            if (hb::launch_init(graph_dev(b, i), b->stream) != hipSuccess) rc = bfail(HB_EHIP, "init kernel failed");
        if (rc != HB_OK) break;
        (void)hipEventRecord(e1, b->stream);
        if (flags & HB_FLAG_GRAPH) {
            if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                rc = bfail(HB_EHIP, "hipStreamBeginCapture failed");
                break;
            }
            const hipError_t ec = launch_iter();
            const hipError_t ee = hipStreamEndCapture(b->stream, &graph);
            if (ec != hipSuccess || ee != hipSuccess || hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
                rc = bfail(HB_EHIP, "hb_batch_run: graph capture failed");
                break;
            }
        }
        std::vector<hb::HbBatchGraph> hs(ng);
        for (;;) {
            const uint32_t k = launches < 8 ? 4u : 16u;
            for (uint32_t i = 0; i < k && rc == HB_OK; ++i) {
                const hipError_t el = gexec ? hipGraphLaunch(gexec, b->stream) : launch_iter();
                if (el != hipSuccess) rc = bfail(HB_EHIP, "iteration launch: %s", hipGetErrorString(el));
            }
            launches += k;
            if (rc != HB_OK) break;
            if (hipMemcpyAsync(hs.data(), b->g, ng * sizeof(hb::HbBatchGraph), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
                hipStreamSynchronize(b->stream) != hipSuccess) {
                rc = bfail(HB_EHIP, "iteration failed: %s", hipGetErrorString(hipGetLastError()));
                break;
            }
            bool all = true;
            for (const auto &g : hs) all = all && g.st.done;
            if (all) break;
        }
        if (rc != HB_OK) break;
        (void)hipEventRecord(e2, b->stream);
        (void)hipEventSynchronize(e2);
        b->h_final = hs;
        for (int i = 0; i < ng; ++i)
            if (hs[i].overflow) rc = bfail(HB_ERANGE, "hb_batch_run: graph %d needs more than %u iterations", i, hs[i].rec_cap - 2);
    } while (0);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    float ms_init = 0.0f, ms_loop = 0.0f;
    if (rc == HB_OK) {
        (void)hipEventElapsedTime(&ms_init, e0, e1);
        (void)hipEventElapsedTime(&ms_loop, e1, e2);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (rc != HB_OK) return rc;
    b->h_rec.resize(b->total_rec);
    HB_BTRY(hipMemcpy(b->h_rec.data(), b->rec, b->total_rec * sizeof(HbIterRec), hipMemcpyDeviceToHost));
    b->ran = true;
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->launches = (int32_t)launches;
        stats->ms_total = ms_loop;
        stats->ms_init = ms_init;
        stats->tiles = b->ntiles;
        for (int i = 0; i < ng; ++i) {
            const auto &g = b->h_final[i];
            stats->iterations_total += g.st.executed;
            stats->iterations_max = std::max<int32_t>(stats->iterations_max, (int32_t)g.st.executed);
            const uint64_t R = 1ull << g.p;
            for (uint32_t t = 1; t <= g.st.executed; ++t) {
                const HbIterRec &r = b->h_rec[g.rec_base + t];
                stats->arcs_gathered += r.arcs[0];
                stats->nodes_recomputed += r.nodes[0];
                stats->alg_bytes += r.nodes[0] * (8 + 2 * R + 40) + r.arcs[0] * (4 + R);
            }
        }
    }
    hb::set_error("");
    return HB_OK;
}

int hb_batch_graph_stats(hb_batch *b, int graph, int32_t *iterations, int32_t *converged, uint64_t *increases,
                         int cap)
{
    if (!b || graph < 0 || graph >= (int)b->hg.size()) return bfail(HB_EINVAL, "hb_batch_graph_stats: bad graph");
    if (!b->ran) return bfail(HB_ESTATE, "hb_batch_graph_stats: no completed run");
    const auto &g = b->h_final[graph];
    if (iterations) *iterations = (int32_t)g.st.executed;
    if (converged) *converged = (int32_t)g.st.converged;
    const int k = std::min<int>(cap, (int)g.st.executed);
    for (int t = 1; t <= k; ++t) increases[t - 1] = b->h_rec[g.rec_base + t].inc[0];
    return k;
}

static const uint8_t *batch_regs(hb_batch *b, int graph, int which)
{
    const auto &g = b->h_final[graph];
    if (which == HB_REGS_REFERENCE) return b->regs0 + g.reg_base;
    return ((g.st.parity ^ 1u) ? b->regs1 : b->regs0) + g.reg_base;
}

int hb_batch_read_registers(hb_batch *b, int graph, int which, uint8_t *dst)
{
    if (!b || !dst || graph < 0 || graph >= (int)b->hg.size() || (which != 0 && which != 1))
        return bfail(HB_EINVAL, "hb_batch_read_registers: bad arguments");
    if (!b->ran) return bfail(HB_ESTATE, "hb_batch_read_registers: no completed run");
    HB_BTRY(hipSetDevice(b->device));
    const auto &g = b->hg[graph];
    HB_BTRY(hipMemcpy(dst, batch_regs(b, graph, which), (size_t)g.n << g.p, hipMemcpyDeviceToHost));
    return HB_OK;
}

int hb_batch_read_estimates(hb_batch *b, int graph, int which, float *dst)
{
    if (!b || !dst || graph < 0 || graph >= (int)b->hg.size() || (which != 0 && which != 1))
        return bfail(HB_EINVAL, "hb_batch_read_estimates: bad arguments");
    if (!b->ran) return bfail(HB_ESTATE, "hb_batch_read_estimates: no completed run");
    HB_BTRY(hipSetDevice(b->device));
    const auto &g = b->hg[graph];
    HB_BTRY(hb::launch_estimates(graph_dev(b, graph), batch_regs(b, graph, which), b->est_tmp, b->stream));
    HB_BTRY(hipMemcpyAsync(dst, b->est_tmp, (size_t)g.n * 4, hipMemcpyDeviceToHost, b->stream));
    HB_BTRY(hipStreamSynchronize(b->stream));
    return HB_OK;
}

int hb_batch_read_accumulators(hb_batch *b, int graph, double *dist, double *inv)
{
    if (!b || graph < 0 || graph >= (int)b->hg.size()) return bfail(HB_EINVAL, "hb_batch_read_accumulators: bad arguments");
    if (!b->ran) return bfail(HB_ESTATE, "hb_batch_read_accumulators: no completed run");
    HB_BTRY(hipSetDevice(b->device));
    const auto &g = b->hg[graph];
    if (dist) HB_BTRY(hipMemcpy(dist, b->dist + g.node_base, (size_t)g.n * 8, hipMemcpyDeviceToHost));
    if (inv) HB_BTRY(hipMemcpy(inv, b->inv + g.node_base, (size_t)g.n * 8, hipMemcpyDeviceToHost));
    return HB_OK;
}

}  // extern "C"

// Coordinate-descent autotuner over the batch tiling knobs (full batch runs).
extern "C" int hb_batch_autotune(hb_batch *b, uint32_t *small_max_deg, uint32_t *mid_max_deg,
                                 uint32_t *tile_arcs, double *trial_ms)
{
    if (!b) return bfail(HB_EINVAL, "hb_batch_autotune: NULL batch");
    hb_batch_stats st;
    auto trial = [&](uint32_t s, uint32_t m, uint32_t t, double *best) -> int {
        int rc = hb_batch_set_schedule(b, s, m, t);
        if (rc != HB_OK) return rc;
        *best = 1e300;
        for (int r = 0; r < 2; ++r) {
            rc = hb_batch_run(b, HB_FLAG_GRAPH, &st);
            if (rc != HB_OK) return rc;
            *best = std::min(*best, st.ms_total);
        }
        return HB_OK;
    };
    uint32_t knob[3] = {b->small_max, b->mid_max, b->tile_arcs};
    const uint32_t cand[3][4] = {{2, 4, 8, 16}, {16, 32, 64, 256}, {2048, 8192, 16384, 65536}};
    double best = 0.0;
    int rc = trial(knob[0], knob[1], knob[2], &best);
    if (rc != HB_OK) return rc;
    for (int k = 0; k < 3; ++k) {
        for (uint32_t v : cand[k]) {
            if (v == knob[k]) continue;
            uint32_t t[3] = {knob[0], knob[1], knob[2]};
            t[k] = v;
            if (t[1] < t[0]) continue;
            double ms = 0.0;
            rc = trial(t[0], t[1], t[2], &ms);
            if (rc != HB_OK) return rc;
            if (ms < best) {
                best = ms;
                knob[k] = v;
            }
        }
    }
    rc = hb_batch_set_schedule(b, knob[0], knob[1], knob[2]);
    if (rc != HB_OK) return rc;
    if (small_max_deg) *small_max_deg = knob[0];
    if (mid_max_deg) *mid_max_deg = knob[1];
    if (tile_arcs) *tile_arcs = knob[2];
    if (trial_ms) *trial_ms = best;
    b->ran = false;
    hb::set_error("");
    return HB_OK;
}
