// hb_batch.hip -- many independent HyperBall runs (graphs of different sizes and
// precisions p = 4..8) advanced together.  One iteration of the whole batch is
//     k_batch_begin -> k_batch_compact -> per precision P (the precisions touch
//     disjoint graphs): k_batch_nodes<P, T> for the three bins, k_batch_chunk<P>,
//     k_batch_merge<P>, the precisions' chains packed by their work into at most four
//     parallel branches, one per hardware queue -> k_batch_finalize
// captured in a hipGraph and replayed until every graph has stopped; a precision whose
// graphs have all stopped is left out of the replay (one captured iteration per set of
// running precisions, kept until the schedule changes).
// The work of one precision P comes in four classes: node tiles of the three degree
// bins (team width G, 4G or 64 lanes per node, G = 2^P / 16) and chunks of hub
// neighbour lists, whose partial unions go to scratch and are merged into their hubs
// by k_batch_merge<P>.  k_batch_compact lists, at the start of every iteration, the
// tiles of each class whose graph is still running (largest first); k_batch_nodes<P,T>
// and k_batch_chunk<P> then walk their class's list, wavefront w taking tiles w,
// w + W, ... (W wavefronts in the grid), and return at once when it is empty, so
// graphs that have stopped cost nothing.  Every graph keeps its
// own iteration state (t, 1/t, parity, frontier mode, stop flag) and applies the
// reference's stopping rule on its own (hyper_ball.rs:833-849); a graph that has
// stopped costs nothing.  The per-node work is union_node, the same function the
// single-graph kernels run, so each graph's registers, estimates, accumulators and
// per-iteration increase counts are its own hb_run's.
//
// Per-iteration counters live in a ring of kRing records per graph; the host copies
// the new ones out at every check (at most kCheck launches apart, fewer than kRing),
// so a run has no record cap.  A graph cannot need more than n + 1 iterations (the
// last one changes nothing), so a graph still running after n + 2 stops the run with
// an error rather than letting the host loop go on.  With SPARKX_HB_TRACE set in the
// environment every check prints a progress line to stderr.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
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
static constexpr int kBatchPMin = 4, kBatchPMax = 8, kBatchNP = kBatchPMax - kBatchPMin + 1;
// side streams for the precisions' chains: one per hardware queue of the process (HIP's
// default is 4); a fifth stream would share a queue with another chain and run after it
static constexpr int kSide = 4;
static constexpr uint32_t kRing = 32;    // per-graph iteration records on the device
static constexpr uint32_t kCheck = 8;    // launches between host checks (< kRing)
static constexpr uint32_t kScrRow = 256; // scratch bytes per hub chunk (2^p, p <= 8)
// work classes of one precision: node bins of team width G, 4G and 64, hub chunks
static constexpr int kClasses = 4, kClassChunk = 3;
static constexpr int kQueues = kBatchNP * kClasses;

struct HbBatchRec {  // counters of one iteration of one graph
    unsigned long long inc, chg, arcs, nodes, gathered;
};

struct HbBatchGraph {
    uint64_t off_base, nbr_base, reg_base, node_base, bits_base;
    uint32_t n, p, n_linked, words;
    uint32_t ring_base;   // first of the graph's kRing records
    uint32_t cap;         // iterations after which a running graph is an error (n + 2)
    uint32_t overflow;    // set when the graph reached cap without stopping
    uint32_t pad;
    float am, five_m;
    double lc_thr;
    HbIterState st;
};

// a tile of one precision: `team` lanes per node over lists[first ..
// first + count) (graph-local ids), or, with team == 0, hub chunk task `first`
struct HbBatchTile {
    uint32_t graph, first, count, team;
};
struct HbBatchTask {  // arcs [beg, end) of node's list (graph-local)
    uint32_t graph, node, beg, end;
};
struct HbBatchHub {   // node's chunks are tasks task0 .. task0 + ntask - 1
    uint32_t graph, node, task0, ntask;
};

struct HbBatchDev {
    const uint32_t *off, *nbr;  // every graph's CSR in its degree order (local ids)
    const uint32_t *orig;       // caller's id of each arc's head (hashed in iteration 1)
    const uint32_t *node_id;    // caller's id of each node (hashed at init)
    uint8_t *regs0, *regs1;
    float *est;
    double *dist, *inv;
    uint32_t *bits0, *bits1;
    HbBatchGraph *g;
    uint32_t ngraphs;
    HbBatchRec *rec;
    const uint32_t *active;     // per class: tiles of running graphs, at class_base[q]
    const uint32_t *active_count;  // [kQueues]
    const uint32_t *lists;
    const HbBatchTask *tasks;
    uint8_t *scratch;           // kScrRow bytes per task
    uint32_t *chunk_flag;       // 1 if the task was recomputed this iteration
    const float *lc;            // LC tables of p = 4..8 back to back (lc_off)
    const HbBiasTable *bias;    // bias windows of p = 4..8
    int frontier;
};

struct HbBatchList {  // the tiles (or hubs) of one (precision, class)
    const HbBatchTile *tiles;
    const HbBatchHub *hubs;
    uint32_t count;
    uint32_t queue;      // class index
    uint32_t base;       // first slot of the class in the active list
};

struct HbBatchClasses {  // every class, for the compaction
    HbBatchList q[kQueues];
};

// x as a wave-uniform value (every lane holds lane 0's copy, in scalar registers):
// values loaded through a per-wave index are uniform, but the compiler cannot prove
// it and would keep every field of a graph's state in vector registers
template <class T>
__device__ __forceinline__ T wave_uniform(const T &x)
{
    static_assert(sizeof(T) % 4 == 0, "whole 32-bit words");
    uint32_t w[sizeof(T) / 4];
    memcpy(w, &x, sizeof(T));
#pragma unroll
    for (uint32_t k = 0; k < sizeof(T) / 4; ++k) w[k] = __builtin_amdgcn_readfirstlane(w[k]);
    T y;
    memcpy(&y, w, sizeof(T));
    return y;
}

__host__ __device__ constexpr uint32_t lc_off(int p) { return p <= 4 ? 0u : lc_off(p - 1) + (1u << (p - 1)) + 1u; }
static constexpr uint32_t kLcTotal = lc_off(kBatchPMax + 1);

__device__ __forceinline__ HbBatchRec *ring_rec(const HbBatchDev &b, const HbBatchGraph &g, uint32_t it)
{
    return b.rec + g.ring_base + it % kRing;
}

// the iteration's view of graph g (as the single-graph kernels build it)
__device__ __forceinline__ HbView batch_view(const HbBatchDev &b, const HbBatchGraph &g, uint32_t *s_first)
{
    const uint32_t it = g.st.iter;
    HbView v;
    v.off = b.off + g.off_base;
    v.nbr = b.nbr + g.nbr_base;
    v.orig = b.orig + g.nbr_base;        // nodes are renumbered by degree per graph; the
    v.node_id = b.node_id + g.node_base; // caller's ids are what every sketch inserts
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
    v.gather_bits = v.use_frontier ? v.bitsR : nullptr;
    v.first = it == 1;       // p <= 8: iteration 1 hashes neighbour ids instead of gathering
    v.scr = s_first;
    return v;
}

template <int P, int T>
__device__ __forceinline__ void batch_nodes(const HbView &v, const HbTables &tb, const uint32_t *list,
                                            uint32_t count, HbCounts &c)
{
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1);
    if (v.gather_bits) {
        for (uint32_t i = lane / T; i < count; i += HB_WAVE / T) union_node<P, T, true>(v, tb, list[i], lane % T, c);
    } else {
        for (uint32_t i = lane / T; i < count; i += HB_WAVE / T) union_node<P, T, false>(v, tb, list[i], lane % T, c);
    }
}

// one hub chunk: partial union of its neighbours into scratch (as k_chunk)
template <int P>
__device__ __forceinline__ void batch_chunk(const HbBatchDev &b, const HbView &v, uint32_t task_id,
                                            const HbBatchTask &task, HbCounts &c)
{
    constexpr uint32_t G = (1u << P) / 16, T = HB_WAVE, S = T / G;
    const uint32_t tl = threadIdx.x & (HB_WAVE - 1), cl = tl % G, slot = tl / G;
    const uint32_t *__restrict__ nb = v.nbr + task.beg;
    const uint32_t *__restrict__ nbh = v.orig + task.beg;  // caller's ids, hashed in iteration 1
    const uint32_t len = task.end - task.beg;
    bool active = true;
    if (v.use_frontier) {
        uint32_t any = 0;
        for (uint32_t j = tl; j < len; j += T) any |= bit_test(v.bitsR, nb[j]);
        active = hb_copy_or<T>(any) != 0;
    }
    if (tl == 0) b.chunk_flag[task_id] = active ? 1u : 0u;
    if (!active) return;
    uint32_t gathered = 0;
    uint4 acc = v.use_frontier
                    ? hb_gather_union<P, S, true>(v.in, v.first, v.scr, v.bitsR, nb, nbh, len, slot, cl,
                                                  make_uint4(0, 0, 0, 0), gathered)
                    : hb_gather_union<P, S, false>(v.in, v.first, v.scr, nullptr, nb, nbh, len, slot, cl,
                                                   make_uint4(0, 0, 0, 0), gathered);
    if (cl == 0) c.gathered += gathered;
#pragma unroll
    for (uint32_t s = G; s < T; s <<= 1) acc = hb_bmax4(acc, hb_shfl_xor4<T>(acc, s));
    if (slot == 0) *(uint4 *)(b.scratch + (uint64_t)task_id * kScrRow + cl * 16) = acc;
    if (tl == 0) c.arcs += len;
}

// the LDS estimator tables of precision P from the batch's per-precision arrays
#define HB_BATCH_TABLES(P, b, name)                                                        \
    __shared__ float name##_lc[(1 << (P)) + 1];                                            \
    __shared__ HbBiasTable name##_bias;                                                    \
    constexpr uint32_t name##_off = lc_off(P); /* folded here: a call would give the kernel a stack */ \
    for (uint32_t i_ = threadIdx.x; i_ <= (1u << (P)); i_ += blockDim.x)                   \
        name##_lc[i_] = (b).lc[name##_off + i_];                                           \
    for (uint32_t i_ = threadIdx.x; i_ < sizeof(HbBiasTable) / 8; i_ += blockDim.x)       \
        ((uint64_t *)&name##_bias)[i_] = ((const uint64_t *)((b).bias + ((P) - kBatchPMin)))[i_]; \
    __syncthreads();

__device__ __forceinline__ void wave_sum_counts(HbCounts &c)
{
#pragma unroll
    for (int s = 1; s < HB_WAVE; s <<= 1) {
        c.inc += __shfl_xor(c.inc, s);
        c.chg += __shfl_xor(c.chg, s);
        c.arcs += __shfl_xor(c.arcs, s);
        c.nodes += __shfl_xor(c.nodes, s);
        c.gathered += __shfl_xor(c.gathered, s);
    }
}

__device__ __forceinline__ void add_counts(HbBatchRec *r, const HbCounts &c)
{
    if (c.inc) atomicAdd(&r->inc, c.inc);
    if (c.chg) atomicAdd(&r->chg, c.chg);
    if (c.arcs) atomicAdd(&r->arcs, c.arcs);
    if (c.nodes) atomicAdd(&r->nodes, c.nodes);
    if (c.gathered) atomicAdd(&r->gathered, c.gathered);
}

// Global wavefront index and count of the grid (wave-uniform).
__device__ __forceinline__ uint32_t wave_id() { return __builtin_amdgcn_readfirstlane(blockIdx.x * (kBatchBlock / HB_WAVE) + threadIdx.x / HB_WAVE); }
__device__ __forceinline__ uint32_t wave_count() { return gridDim.x * (kBatchBlock / HB_WAVE); }

// Union over the active node tiles of one (precision P, team width T) class.  Every
// loop has a trip count fixed by the active list and its tiles (at most 4096 nodes).
// Six wavefronts per SIMD from p = 6 up (80 VGPRs: the one-copy bins spill 2-5 where
// seven spilled 9-19): with the precisions' chains overlapping, the p = 7 chain is the
// longest, and the 64-graph run takes 14.1 ms against 14.6 ms at seven.
template <int P, int T>
__global__ __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(P >= 6 ? 6 : 8))) void k_batch_nodes(
    HbBatchDev b, HbBatchList l)
{
    const uint32_t count = b.active_count[l.queue];
    if (count == 0) return;  // the whole grid leaves: no graph of this class is running
    HB_BATCH_TABLES(P, b, tbl)
    __shared__ uint32_t s_first[kBatchBlock * 16];
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1);
    for (uint32_t w = wave_id(); w < count; w += wave_count()) {
        const HbBatchTile tile = wave_uniform(l.tiles[b.active[l.base + w]]);
        const HbBatchGraph g = wave_uniform(b.g[tile.graph]);
        const HbTables tb = {tbl_lc, &tbl_bias, g.am, g.five_m, g.lc_thr};
        const HbView v = batch_view(b, g, s_first);
        HbCounts c;
        batch_nodes<P, T>(v, tb, b.lists + tile.first, tile.count, c);
        wave_sum_counts(c);
        if (lane == 0) add_counts(ring_rec(b, g, g.st.iter), c);
    }
}

// Partial unions of the active hub chunks of precision P, one wavefront per chunk.
template <int P>
__global__ __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(P >= 7 ? 7 : 8))) void k_batch_chunk(
    HbBatchDev b, HbBatchList l)
{
    const uint32_t count = b.active_count[l.queue];
    if (count == 0) return;
    __shared__ uint32_t s_first[kBatchBlock * 16];
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1);
    for (uint32_t w = wave_id(); w < count; w += wave_count()) {
        const HbBatchTile tile = wave_uniform(l.tiles[b.active[l.base + w]]);
        const HbBatchGraph g = wave_uniform(b.g[tile.graph]);
        const HbView v = batch_view(b, g, s_first);
        HbCounts c;
        batch_chunk<P>(b, v, tile.first, wave_uniform(b.tasks[tile.first]), c);
        wave_sum_counts(c);
        if (lane == 0) add_counts(ring_rec(b, g, g.st.iter), c);
    }
}

// Start of an iteration, one block per class: the tiles of graphs still running, in
// list order (largest first), into the class's active list.
__global__ __launch_bounds__(1024) void k_batch_compact(HbBatchDev b, HbBatchClasses cl)
{
    __shared__ uint32_t s_wave[1024 / HB_WAVE];
    __shared__ uint32_t s_total;
    const HbBatchList l = cl.q[blockIdx.x];
    uint32_t *out = (uint32_t *)b.active + l.base;
    const uint32_t lane = threadIdx.x & (HB_WAVE - 1), wv = threadIdx.x / HB_WAVE;
    if (threadIdx.x == 0) s_total = 0;
    __syncthreads();
    for (uint32_t base = 0; base < l.count; base += 1024) {  // uniform trip count
        const uint32_t i = base + threadIdx.x;
        const bool keep = i < l.count && !b.g[l.tiles[i].graph].st.done;
        const unsigned long long m = __ballot(keep);
        if (lane == 0) s_wave[wv] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = s_total;
        for (uint32_t k = 0; k < wv; ++k) before += s_wave[k];
        if (keep) out[before + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = i;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (uint32_t k = 0; k < 1024 / HB_WAVE; ++k) t += s_wave[k];
            s_total += t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) ((uint32_t *)b.active_count)[blockIdx.x] = s_total;
}

// hubs of precision P: one wavefront per hub merges its own registers with the
// partials of its recomputed chunks (as k_merge)
template <int P>
__global__ __launch_bounds__(kBatchBlock) void k_batch_merge(HbBatchDev b, HbBatchList l)
{
    constexpr uint32_t G = (1u << P) / 16, T = HB_WAVE, S = T / G;
    HB_BATCH_TABLES(P, b, tbl)
    const uint32_t tl = threadIdx.x & (HB_WAVE - 1), cl = tl % G, slot = tl / G;
    for (uint32_t h = wave_id(); h < l.count; h += wave_count()) {
        const HbBatchHub hub = wave_uniform(l.hubs[h]);
        const HbBatchGraph g = wave_uniform(b.g[hub.graph]);
        if (g.st.done) continue;
        const HbTables tb = {tbl_lc, &tbl_bias, g.am, g.five_m, g.lc_thr};
        const HbView v = batch_view(b, g, nullptr);
        const uint32_t u = hub.node;
        HbCounts c;
        uint32_t any = 0;
        for (uint32_t j = tl; j < hub.ntask; j += T) any |= b.chunk_flag[hub.task0 + j];
        if (hb_copy_or<T>(any) == 0) {
            // no chunk saw a changed neighbour: c_t(u) = c_{t-1}(u)
            if (slot == 0 && bit_test(v.bitsR, u))
                *(uint4 *)(v.out + ((uint64_t)u << P) + cl * 16) = *(const uint4 *)(v.in + ((uint64_t)u << P) + cl * 16);
            continue;
        }
        const uint4 own = v.first ? hb_own_from_word(hb_insertion<P>(v.node_id[u]), cl)
                                  : *(const uint4 *)(v.in + ((uint64_t)u << P) + cl * 16);
        uint4 acc = own;
        for (uint32_t j = slot; j < hub.ntask; j += S)
            if (b.chunk_flag[hub.task0 + j])
                acc = hb_bmax4(acc, *(const uint4 *)(b.scratch + (uint64_t)(hub.task0 + j) * kScrRow + cl * 16));
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
            c.nodes += 1;
            add_counts(ring_rec(b, g, g.st.iter), c);
        }
    }
}

// start of an iteration: for every running graph clear the bitmap this iteration
// writes and its ring record
__global__ __launch_bounds__(kBatchBlock) void k_batch_begin(HbBatchDev b)
{
    for (uint32_t gi = blockIdx.x; gi < b.ngraphs; gi += gridDim.x) {
        const HbBatchGraph &g = b.g[gi];
        if (g.st.done) continue;
        uint32_t *bw = ((g.st.iter & 1u) ? b.bits1 : b.bits0) + g.bits_base;
        for (uint32_t i = threadIdx.x; i < g.words; i += blockDim.x) bw[i] = 0u;
        if (threadIdx.x == 0) *ring_rec(b, g, g.st.iter) = HbBatchRec{0, 0, 0, 0, 0};
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
    const HbBatchRec &r = *ring_rec(b, g, it);
    const bool truncated = st.max_iters && it >= st.max_iters;
    if (r.inc == 0 || truncated || it >= g.cap) {
        st.done = 1;
        st.executed = it;
        st.converged = r.inc == 0;
        g.overflow = r.inc != 0 && !truncated;
        return;
    }
    st.iter = it + 1;
    st.parity ^= 1u;
    st.t = st.t + 1.0;
    st.inv_t = __ddiv_rn(1.0, st.t);
    st.skip_scan = 2 * r.chg > (unsigned long long)g.n_linked;
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
struct DevArr {
    T *ptr = nullptr;
    size_t count = 0;
    hipError_t alloc(size_t n)
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = n;
        return hipMalloc(&ptr, std::max<size_t>(n, 1) * sizeof(T));
    }
    hipError_t put(const std::vector<T> &h, hipStream_t s)
    {
        const hipError_t e = alloc(h.size());
        if (e != hipSuccess || h.empty()) return e;
        return hipMemcpyAsync(ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s);
    }
    ~DevArr()
    {
        if (ptr) (void)hipFree(ptr);
    }
};

bool trace_on()
{
    const char *e = std::getenv("SPARKX_HB_TRACE");
    return e && *e && strcmp(e, "0") != 0;
}

}  // namespace

struct hb_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    // the precisions touch disjoint graphs, so their kernels run as up to kSide chains on
    // streams of their own between a fork after the tile compaction and a join before
    // finalize (in the captured graph: parallel branches)
    hipStream_t side[hb::kSide] = {};
    hipEvent_t fork = nullptr, join[hb::kSide] = {};
    // captured iterations, one per (precisions still running, dense flag); they depend
    // only on the batch's buffers and tiles, so they are kept until the tiles change
    std::vector<std::pair<uint32_t, hipGraphExec_t>> execs;
    void drop_execs()
    {
        for (const auto &x : execs) (void)hipGraphExecDestroy(x.second);
        execs.clear();
    }
    std::vector<hb::HbBatchGraph> hg;  // host copy (bases, constants)
    std::vector<std::vector<uint32_t>> h_off;
    std::vector<int> max_iters;
    uint64_t total_nodes = 0, total_arcs = 0, total_regs = 0, total_words = 0;
    // CSRs in each graph's degree order, the caller's ids (orig per arc, perm per node)
    // and the inverse numbering (sigma: caller's node -> device node)
    DevArr<uint32_t> off, nbr, orig, perm, sigma, bits0, bits1, lists, chunk_flag;
    DevArr<uint32_t> active, active_count;  // per class: tiles of running graphs
    DevArr<uint8_t> regs0, regs1, scratch, readback;
    DevArr<float> est, lc_all, est_tmp;
    DevArr<double> dist, inv;
    DevArr<HbBiasTable> bias_all;
    DevArr<hb::HbBatchGraph> g;
    DevArr<hb::HbBatchRec> rec;
    DevArr<hb::HbBatchTile> tiles[hb::kQueues];  // per (precision, class)
    DevArr<hb::HbBatchHub> hubs[hb::kBatchNP];
    DevArr<hb::HbBatchTask> tasks;
    float *lc_p[hb::kBatchPMax + 1] = {};  // device pointers into lc_all, per p
    uint32_t ntiles[hb::kQueues] = {}, tile_base[hb::kQueues] = {}, nhubs[hb::kBatchNP] = {}, ntasks = 0;
    unsigned grid[hb::kQueues] = {};
    // 0: per-precision defaults.  A tile should be a few rounds of one wavefront: a
    // class with fewer tiles than resident wavefronts (7,168) leaves most of them idle
    // while a few walk long tiles node by node.
    uint32_t small_max = 0, mid_max = 0, tile_arcs = 1024;
    bool sched_ok = false, ran = false;
    std::vector<hb::HbBatchGraph> h_final;
    std::vector<std::vector<hb::HbBatchRec>> history;  // per graph, iteration t at [t - 1]

    ~hb_batch()
    {
        drop_execs();
        for (int i = 0; i < hb::kSide; ++i) {
            if (join[i]) (void)hipEventDestroy(join[i]);
            if (side[i]) (void)hipStreamDestroy(side[i]);
        }
        if (fork) (void)hipEventDestroy(fork);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static HbDev graph_dev(hb_batch *b, int gi)
{
    const hb::HbBatchGraph &g = b->hg[gi];
    HbDev d{};
    d.off = b->off.ptr + g.off_base;
    d.nbr = b->nbr.ptr + g.nbr_base;
    d.orig = b->orig.ptr + g.nbr_base;
    d.node_id = b->perm.ptr + g.node_base;
    d.regs0 = b->regs0.ptr + g.reg_base;
    d.regs1 = b->regs1.ptr + g.reg_base;
    d.est = b->est.ptr + g.node_base;
    d.dist = b->dist.ptr + g.node_base;
    d.inv = b->inv.ptr + g.node_base;
    d.lc = b->lc_p[g.p];
    d.bias = b->bias_all.ptr + (g.p - hb::kBatchPMin);
    d.n = g.n;
    d.n_linked = g.n_linked;
    d.p = g.p;
    d.am = g.am;
    d.five_m = g.five_m;
    d.lc_thr = g.lc_thr;
    return d;
}

// degree bounds of one graph's bins (the single-graph engine's defaults unless set)
static void bin_bounds(const hb_batch *b, uint32_t p, uint32_t *small_max, uint32_t *mid_max, uint32_t *wave_max,
                       uint32_t *chunk)
{
    *chunk = 2048;
    *small_max = b->small_max ? b->small_max : (p <= 5 ? 64u : 256u);
    *wave_max = p <= 5 ? 16384u : 2048u;
    *wave_max = std::max(*wave_max, *small_max);
    *mid_max = b->mid_max ? std::max(b->mid_max, *small_max) : std::max(*small_max, std::min(*wave_max, 8 * *small_max));
    *mid_max = std::min(*mid_max, *wave_max);
}

static int build_tiles(hb_batch *b)
{
    if (b->sched_ok) return HB_OK;
    b->drop_execs();
    struct Item {
        hb::HbBatchTile t;
        uint64_t work;
    };
    std::vector<uint32_t> lists;
    std::vector<hb::HbBatchTask> tasks;
    std::vector<Item> items[hb::kQueues];
    std::vector<hb::HbBatchHub> hubs[hb::kBatchNP];
    for (size_t gi = 0; gi < b->hg.size(); ++gi) {
        const hb::HbBatchGraph &hgi = b->hg[gi];
        const auto &off = b->h_off[gi];
        const uint32_t p = hgi.p, G = (1u << p) / 16u, pi = p - hb::kBatchPMin;
        const uint32_t teams[3] = {G, std::min(64u, 4 * G), 64u};
        uint32_t small_max, mid_max, wave_max, chunk;
        bin_bounds(b, p, &small_max, &mid_max, &wave_max, &chunk);
        std::vector<uint32_t> bins[3];
        for (uint32_t u = 0; u < hgi.n; ++u) {
            const uint32_t d = off[u + 1] - off[u];
            if (d == 0) continue;
            if (d <= small_max) bins[0].push_back(u);
            else if (d <= mid_max) bins[1].push_back(u);
            else if (d <= wave_max) bins[2].push_back(u);
            else {  // hub: chunks of `chunk` arcs, each one tile of its own
                hb::HbBatchHub h{(uint32_t)gi, u, (uint32_t)tasks.size(), 0};
                for (uint32_t e = off[u]; e < off[u + 1]; e += chunk, ++h.ntask) {
                    const uint32_t end = std::min(e + chunk, off[u + 1]);
                    items[pi * hb::kClasses + hb::kClassChunk].push_back(
                        {{(uint32_t)gi, (uint32_t)tasks.size(), 1, 0}, (uint64_t)(end - e + 8) << p});
                    tasks.push_back({(uint32_t)gi, u, e, end});
                }
                hubs[pi].push_back(h);
            }
        }
        for (int k = 0; k < 3; ++k) {
            const uint32_t per_wave = 64u / teams[k];
            size_t i = 0;
            while (i < bins[k].size()) {
                Item it{};
                it.t.graph = (uint32_t)gi;
                it.t.first = (uint32_t)lists.size();
                it.t.team = teams[k];
                uint64_t arcs = 0;
                uint32_t cnt = 0;
                // a tile holds whole waves' worth of nodes until it has tile_arcs arcs
                while (i < bins[k].size() && (arcs < b->tile_arcs || cnt % per_wave) && cnt < 4096) {
                    const uint32_t u = bins[k][i++];
                    lists.push_back(u);
                    arcs += off[u + 1] - off[u] + 8;
                    ++cnt;
                }
                it.t.count = cnt;
                it.work = arcs << p;
                items[pi * hb::kClasses + k].push_back(it);
            }
        }
    }
    HB_BTRY(b->lists.put(lists, b->stream));
    HB_BTRY(b->tasks.put(tasks, b->stream));
    b->ntasks = (uint32_t)tasks.size();
    HB_BTRY(b->scratch.alloc((size_t)b->ntasks * hb::kScrRow));
    HB_BTRY(b->chunk_flag.alloc(b->ntasks));
    int dev_cus = 0;
    HB_BTRY(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, b->device));
    for (int q = 0; q < hb::kQueues; ++q) {
        // largest first: wavefront w takes active tiles w, w + W, ..., so the longest
        // ones start first and the short ones fill in behind them
        std::stable_sort(items[q].begin(), items[q].end(), [](const Item &x, const Item &y) { return x.work > y.work; });
        std::vector<hb::HbBatchTile> t(items[q].size());
        for (size_t i = 0; i < t.size(); ++i) t[i] = items[q][i].t;
        HB_BTRY(b->tiles[q].put(t, b->stream));
        b->ntiles[q] = (uint32_t)t.size();
        // persistent grid: as many blocks as fit at once (7 of 4 wavefronts per CU at
        // 7 wavefronts per SIMD), no more than there are tiles
        b->grid[q] = (unsigned)std::min<uint64_t>(std::max<uint32_t>(b->ntiles[q], 1), 7ull * (unsigned)std::max(dev_cus, 1));
    }
    for (int pi = 0; pi < hb::kBatchNP; ++pi) {
        HB_BTRY(b->hubs[pi].put(hubs[pi], b->stream));
        b->nhubs[pi] = (uint32_t)hubs[pi].size();
    }
    uint32_t all = 0;
    for (int q = 0; q < hb::kQueues; ++q) {
        b->tile_base[q] = all;
        all += b->ntiles[q];
    }
    HB_BTRY(b->active.alloc(all));
    HB_BTRY(hipStreamSynchronize(b->stream));
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
    HB_BTRY(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
    for (int i = 0; i < hb::kSide; ++i) {
        HB_BTRY(hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking));
        HB_BTRY(hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming));
    }
    b->hg.resize(count);
    b->h_off.resize(count);
    b->max_iters.resize(count);
    uint64_t nodes = 0, arcs = 0, regs = 0, words = 0;
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
        g.ring_base = (uint32_t)i * hb::kRing;
        g.n = (uint32_t)gd.n;
        g.p = (uint32_t)gd.p;
        g.words = (uint32_t)((gd.n + 31) / 32);
        g.cap = (uint32_t)gd.n + 2;
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
    }
    b->total_nodes = nodes;
    b->total_arcs = arcs;
    b->total_regs = regs;
    b->total_words = words;
    HB_BTRY(b->off.alloc(nodes + count));
    HB_BTRY(b->nbr.alloc(arcs));
    HB_BTRY(b->regs0.alloc(regs));
    HB_BTRY(b->regs1.alloc(regs));
    HB_BTRY(b->est.alloc(nodes));
    HB_BTRY(b->est_tmp.alloc(nodes));
    HB_BTRY(b->dist.alloc(nodes));
    HB_BTRY(b->inv.alloc(nodes));
    HB_BTRY(b->bits0.alloc(words));
    HB_BTRY(b->bits1.alloc(words));
    HB_BTRY(b->active_count.alloc(hb::kQueues));
    HB_BTRY(b->g.alloc(count));
    HB_BTRY(b->rec.alloc((size_t)count * hb::kRing));
    // estimator tables of p = 4..8: LC values (f64 then f32, as m ln(m/z)) and bias windows
    std::vector<float> lc(hb::kLcTotal, 0.0f);
    for (int p = hb::kBatchPMin; p <= hb::kBatchPMax; ++p) {
        const uint32_t m = 1u << p;
        for (uint32_t z = 1; z <= m; ++z) lc[hb::lc_off(p) + z] = (float)((double)m * std::log((double)m / (double)z));
    }
    HB_BTRY(b->lc_all.alloc(hb::kLcTotal));
    HB_BTRY(hipMemcpy(b->lc_all.ptr, lc.data(), lc.size() * 4, hipMemcpyHostToDevice));
    for (int p = hb::kBatchPMin; p <= hb::kBatchPMax; ++p) b->lc_p[p] = b->lc_all.ptr + hb::lc_off(p);
    HB_BTRY(b->bias_all.alloc(hb::kBatchNP));
    for (int p = hb::kBatchPMin; p <= hb::kBatchPMax; ++p)
        HB_BTRY(hipMemcpy(b->bias_all.ptr + (p - hb::kBatchPMin), hb::bias_table(p), sizeof(HbBiasTable),
                          hipMemcpyHostToDevice));
    // CSR upload (per graph, converted to u32)
    std::vector<uint32_t> tmp;
    for (int i = 0; i < count; ++i) {
        const hb_graph_desc &gd = graphs[i];
        const hb::HbBatchGraph &g = b->hg[i];
        HB_BTRY(hipMemcpy(b->off.ptr + g.off_base, b->h_off[i].data(), (gd.n + 1) * 4, hipMemcpyHostToDevice));
        tmp.resize(gd.arcs);
        for (uint64_t k = 0; k < gd.arcs; ++k)
            tmp[k] = (uint32_t)(gd.nbr_bytes == 8 ? ((const uint64_t *)gd.nbrs)[k] : ((const uint32_t *)gd.nbrs)[k]);
        if (gd.arcs) HB_BTRY(hipMemcpy(b->nbr.ptr + g.nbr_base, tmp.data(), gd.arcs * 4, hipMemcpyHostToDevice));
    }
    // every graph renumbered by descending degree on the device (as hb_run's default
    // layout): its most gathered register rows sit together; the tiles are built from
    // the renumbered offsets
    {
        DevArr<uint32_t> roff, rnbr;
        HB_BTRY(roff.alloc(nodes + count));
        HB_BTRY(rnbr.alloc(arcs));
        HB_BTRY(b->orig.alloc(arcs));
        HB_BTRY(b->perm.alloc(nodes));
        HB_BTRY(b->sigma.alloc(nodes));
        for (int i = 0; i < count; ++i) {
            const hb::HbBatchGraph &g = b->hg[i];
            HB_BTRY(hb::relabel_by_degree(b->off.ptr + g.off_base, b->nbr.ptr + g.nbr_base, g.n,
                                          (uint32_t)graphs[i].arcs, b->perm.ptr + g.node_base,
                                          b->sigma.ptr + g.node_base, roff.ptr + g.off_base, rnbr.ptr + g.nbr_base,
                                          b->orig.ptr + g.nbr_base, b->stream));
            HB_BTRY(hipMemcpy(b->h_off[i].data(), roff.ptr + g.off_base, ((size_t)g.n + 1) * 4, hipMemcpyDeviceToHost));
        }
        std::swap(b->off.ptr, roff.ptr);
        std::swap(b->nbr.ptr, rnbr.ptr);
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
    if (mid_max_deg) b->mid_max = mid_max_deg;
    if (tile_arcs) b->tile_arcs = tile_arcs;
    b->sched_ok = false;
    return HB_OK;
}

int hb_batch_run(hb_batch *b, uint32_t flags, hb_batch_stats *stats)
{
    if (!b) return bfail(HB_EINVAL, "hb_batch_run: NULL batch");
    HB_BTRY(hipSetDevice(b->device));
    {
        const int rc = build_tiles(b);
        if (rc != HB_OK) return rc;
    }
    b->ran = false;
    const int ng = (int)b->hg.size();
    std::vector<hb::HbBatchGraph> init = b->hg;
    uint32_t cap_max = 0;
    for (int i = 0; i < ng; ++i) {
        HbIterState &st = init[i].st;
        st = HbIterState{};
        st.iter = 1;
        st.max_iters = (uint32_t)b->max_iters[i];
        st.t = 1.0;
        st.inv_t = 1.0;
        init[i].overflow = 0;
        cap_max = std::max(cap_max, st.max_iters ? std::min(st.max_iters, init[i].cap) : init[i].cap);
    }
    hb::HbBatchDev d{};
    d.off = b->off.ptr;
    d.nbr = b->nbr.ptr;
    d.orig = b->orig.ptr;
    d.node_id = b->perm.ptr;
    d.regs0 = b->regs0.ptr;
    d.regs1 = b->regs1.ptr;
    d.est = b->est.ptr;
    d.dist = b->dist.ptr;
    d.inv = b->inv.ptr;
    d.bits0 = b->bits0.ptr;
    d.bits1 = b->bits1.ptr;
    d.g = b->g.ptr;
    d.ngraphs = (uint32_t)ng;
    d.rec = b->rec.ptr;
    d.active = b->active.ptr;
    d.active_count = b->active_count.ptr;
    d.lists = b->lists.ptr;
    d.tasks = b->tasks.ptr;
    d.scratch = b->scratch.ptr;
    d.chunk_flag = b->chunk_flag.ptr;
    d.lc = b->lc_all.ptr;
    d.bias = b->bias_all.ptr;
    d.frontier = (flags & HB_FLAG_DENSE) ? 0 : 1;
    hb::HbBatchList lists[hb::kQueues], hub_lists[hb::kBatchNP];
    hb::HbBatchClasses classes;
    for (int q = 0; q < hb::kQueues; ++q)
        classes.q[q] = lists[q] = {b->tiles[q].ptr, nullptr, b->ntiles[q], (uint32_t)q, b->tile_base[q]};
    for (int pi = 0; pi < hb::kBatchNP; ++pi) hub_lists[pi] = {nullptr, b->hubs[pi].ptr, b->nhubs[pi], 0, 0};

    hipEvent_t e0, e1, e2;
    HB_BTRY(hipEventCreate(&e0));
    HB_BTRY(hipEventCreate(&e1));
    HB_BTRY(hipEventCreate(&e2));
    int rc = HB_OK;
    uint32_t launches = 0;
    const bool trace = trace_on();
    // precisions with any tile or hub: each gets a branch
    bool has_work[hb::kBatchNP] = {};
    for (int pi = 0; pi < hb::kBatchNP; ++pi) {
        has_work[pi] = hub_lists[pi].count != 0;
        for (int c = 0; c < hb::kClasses; ++c) has_work[pi] = has_work[pi] || lists[pi * hb::kClasses + c].count;
    }
    // bytes of rows a dense iteration gathers per precision, to balance the side streams
    uint64_t pwork[hb::kBatchNP] = {};
    for (int i = 0; i < ng; ++i)
        pwork[b->hg[i].p - hb::kBatchPMin] += ((uint64_t)b->h_off[i].back() + b->hg[i].n) << b->hg[i].p;
    // `running`: bit pi set while a graph of precision kBatchPMin + pi may still run.  A
    // stopped graph never runs again and its kernels would return at once, so leaving
    // out a precision whose graphs have all stopped changes nothing.
    auto launch_iter = [&](uint32_t running) -> hipError_t {
        hb::k_batch_begin<<<(unsigned)std::min(ng, 1024), hb::kBatchBlock, 0, b->stream>>>(d);
        hb::k_batch_compact<<<hb::kQueues, 1024, 0, b->stream>>>(d, classes);
        hipError_t e = hipEventRecord(b->fork, b->stream);
        // the running precisions' chains (the three node bins of team widths G, 4G capped at
        // 64, and 64, the hub chunks, then the hubs), heaviest first, each on the side
        // stream with the least work so far
        int order[hb::kBatchNP], side_of[hb::kBatchNP], np = 0;
        for (int pi = 0; pi < hb::kBatchNP; ++pi)
            if (has_work[pi] && ((running >> pi) & 1u)) order[np++] = pi;
        std::stable_sort(order, order + np, [&](int x, int y) { return pwork[x] > pwork[y]; });
        uint64_t load[hb::kSide] = {};
        for (int k = 0; k < np; ++k) {
            const int si = (int)(std::min_element(load, load + hb::kSide) - load);
            side_of[order[k]] = si;
            load[si] += pwork[order[k]];
        }
#define HB_BATCH_LAUNCH(q, kernel)                                                                     \
        if (lists[q].count) kernel<<<b->grid[q], hb::kBatchBlock, 0, s>>>(d, lists[q]);
#define HB_BATCH_P(P)                                                                                  \
        case P: {                                                                                      \
            constexpr int G_ = (1 << P) / 16, q0 = (P - hb::kBatchPMin) * hb::kClasses;               \
            HB_BATCH_LAUNCH(q0, (hb::k_batch_nodes<P, G_>))                                            \
            HB_BATCH_LAUNCH(q0 + 1, (hb::k_batch_nodes<P, (4 * G_ < 64 ? 4 * G_ : 64)>))              \
            HB_BATCH_LAUNCH(q0 + 2, (hb::k_batch_nodes<P, 64>))                                        \
            HB_BATCH_LAUNCH(q0 + hb::kClassChunk, hb::k_batch_chunk<P>)                                \
            if (hub_lists[P - hb::kBatchPMin].count)                                                   \
                hb::k_batch_merge<P><<<(hub_lists[P - hb::kBatchPMin].count + 3) / 4, hb::kBatchBlock, 0, s>>>( \
                    d, hub_lists[P - hb::kBatchPMin]);                                                 \
        } break;
        for (int si = 0; si < hb::kSide && e == hipSuccess; ++si) {
            const hipStream_t s = b->side[si];
            bool forked = false;
            for (int k = 0; k < np && e == hipSuccess; ++k) {
                if (side_of[order[k]] != si) continue;
                if (!forked) {
                    e = hipStreamWaitEvent(s, b->fork, 0);
                    forked = true;
                }
                switch (order[k] + hb::kBatchPMin) { HB_BATCH_P(4) HB_BATCH_P(5) HB_BATCH_P(6) HB_BATCH_P(7) HB_BATCH_P(8) }
            }
            if (forked && e == hipSuccess) e = hipEventRecord(b->join[si], s);
            if (forked && e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join[si], 0);
        }
#undef HB_BATCH_P
#undef HB_BATCH_LAUNCH
        hb::k_batch_finalize<<<(ng + 63) / 64, 64, 0, b->stream>>>(d);
        const hipError_t el = hipGetLastError();
        return e != hipSuccess ? e : el;
    };
    b->history.assign(ng, {});
    const auto t_start = std::chrono::steady_clock::now();
    do {
        if (hipMemcpyAsync(b->g.ptr, init.data(), ng * sizeof(hb::HbBatchGraph), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
            hb::launch_fill(b->bits0.ptr, b->total_words, 0u, b->stream) != hipSuccess ||
            hb::launch_fill(b->bits1.ptr, b->total_words, 0u, b->stream) != hipSuccess ||
            hipEventRecord(e0, b->stream) != hipSuccess) {
            rc = bfail(HB_EHIP, "hb_batch_run: set-up failed");
            break;
        }
        for (int i = 0; i < ng && rc == HB_OK; ++i)
            if (hb::launch_init(graph_dev(b, i), b->stream) != hipSuccess) rc = bfail(HB_EHIP, "init kernel failed");
        if (rc != HB_OK) break;
        (void)hipEventRecord(e1, b->stream);
        // the replayed iteration for a set of precisions, captured on first use
        auto exec_for = [&](uint32_t running, hipGraphExec_t *out) -> int {
            const uint32_t key = running | (d.frontier ? 0u : 1u << 16);
            for (const auto &x : b->execs)
                if (x.first == key) {
                    *out = x.second;
                    return HB_OK;
                }
            hipGraph_t graph = nullptr;
            hipGraphExec_t ge = nullptr;
            if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
                return bfail(HB_EHIP, "hipStreamBeginCapture failed");
            const hipError_t ec = launch_iter(running);
            const hipError_t ee = hipStreamEndCapture(b->stream, &graph);
            const bool ok = ec == hipSuccess && ee == hipSuccess &&
                            hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0) == hipSuccess;
            if (graph) (void)hipGraphDestroy(graph);
            if (!ok) return bfail(HB_EHIP, "hb_batch_run: graph capture failed");
            b->execs.emplace_back(key, ge);
            *out = ge;
            return HB_OK;
        };
        uint32_t running_p = (1u << hb::kBatchNP) - 1;
        std::vector<hb::HbBatchGraph> hs(ng);
        std::vector<hb::HbBatchRec> ring((size_t)ng * hb::kRing);
        std::vector<uint32_t> harvested(ng, 0);  // iterations copied into history
        for (;;) {
            // no graph can need more than cap_max iterations; launches past that only
            // happen if a graph failed to stop, which finalize reports as overflow
            const uint32_t k = launches < 8 ? 2u : hb::kCheck;
            if (launches >= cap_max + 2 * hb::kCheck) {
                rc = bfail(HB_ESTATE, "hb_batch_run: %u launches without every graph stopping", launches);
                break;
            }
            hipGraphExec_t gexec = nullptr;
            if (flags & HB_FLAG_GRAPH) rc = exec_for(running_p, &gexec);
            for (uint32_t i = 0; i < k && rc == HB_OK; ++i) {
                const hipError_t el = gexec ? hipGraphLaunch(gexec, b->stream) : launch_iter(running_p);
                if (el != hipSuccess) rc = bfail(HB_EHIP, "iteration launch: %s", hipGetErrorString(el));
            }
            launches += k;
            if (rc != HB_OK) break;
            if (hipMemcpyAsync(hs.data(), b->g.ptr, ng * sizeof(hb::HbBatchGraph), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
                hipMemcpyAsync(ring.data(), b->rec.ptr, ring.size() * sizeof(hb::HbBatchRec), hipMemcpyDeviceToHost,
                               b->stream) != hipSuccess ||
                hipStreamSynchronize(b->stream) != hipSuccess) {
                rc = bfail(HB_EHIP, "iteration failed: %s", hipGetErrorString(hipGetLastError()));
                break;
            }
            bool all = true;
            uint32_t running = 0, it_max = 0;
            running_p = 0;
            for (int i = 0; i < ng; ++i) {
                const hb::HbBatchGraph &g = hs[i];
                if (!g.st.done) running_p |= 1u << (g.p - hb::kBatchPMin);
                const uint32_t complete = g.st.done ? g.st.executed : g.st.iter - 1;  // iterations finished
                for (uint32_t t = harvested[i] + 1; t <= complete; ++t)
                    b->history[i].push_back(ring[(size_t)i * hb::kRing + t % hb::kRing]);
                harvested[i] = complete;
                all = all && g.st.done;
                running += !g.st.done;
                it_max = std::max(it_max, complete);
            }
            if (trace) {
                const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
                fprintf(stderr, "[hb_batch] %.3fs launches %u: %u of %d graphs running, longest at %u iterations\n", s,
                        launches, running, ng, it_max);
                fflush(stderr);
            }
            if (all) break;
        }
        if (rc != HB_OK) break;
        (void)hipEventRecord(e2, b->stream);
        (void)hipEventSynchronize(e2);
        b->h_final = hs;
        for (int i = 0; i < ng; ++i)
            if (hs[i].overflow)
                rc = bfail(HB_ESTATE, "hb_batch_run: graph %d still changing after %u iterations (n + 2)", i, hs[i].cap);
    } while (0);
    float ms_init = 0.0f, ms_loop = 0.0f;
    if (rc == HB_OK) {
        (void)hipEventElapsedTime(&ms_init, e0, e1);
        (void)hipEventElapsedTime(&ms_loop, e1, e2);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (rc != HB_OK) {
        (void)hipStreamSynchronize(b->stream);
        return rc;
    }
    b->ran = true;
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->launches = (int32_t)launches;
        stats->ms_total = ms_loop;
        stats->ms_init = ms_init;
        for (int q = 0; q < hb::kQueues; ++q) stats->tiles += b->ntiles[q];
        for (int i = 0; i < ng; ++i) {
            const auto &g = b->h_final[i];
            stats->iterations_total += g.st.executed;
            stats->iterations_max = std::max<int32_t>(stats->iterations_max, (int32_t)g.st.executed);
            const uint64_t R = 1ull << g.p;
            for (const hb::HbBatchRec &r : b->history[i]) {
                stats->arcs_gathered += r.arcs;
                stats->nodes_recomputed += r.nodes;
                stats->alg_bytes += r.nodes * (8 + 2 * R + 40) + r.arcs * 4 + r.gathered * R;
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
    const int k = std::min<int>(cap, (int)b->history[graph].size());
    if (increases)
        for (int t = 0; t < k; ++t) increases[t] = b->history[graph][t].inc;
    return increases ? k : 0;
}

static const uint8_t *batch_regs(hb_batch *b, int graph, int which)
{
    const auto &g = b->h_final[graph];
    if (which == HB_REGS_REFERENCE) return b->regs0.ptr + g.reg_base;
    // the last iteration read regs[parity] and wrote the other buffer
    return (g.st.parity ? b->regs0.ptr : b->regs1.ptr) + g.reg_base;
}

// n rows of row_bytes of graph `graph` from device src (its device order) to host dst
// in the caller's node order
static int batch_copy_out(hb_batch *b, int graph, const void *src, void *dst, size_t row_bytes)
{
    const auto &g = b->hg[graph];
    const size_t bytes = (size_t)g.n * row_bytes;
    if (b->readback.count < bytes) HB_BTRY(b->readback.alloc(bytes));
    HB_BTRY(hb::unpermute_rows(src, b->readback.ptr, b->sigma.ptr + g.node_base, g.n, row_bytes, b->stream));
    HB_BTRY(hipMemcpyAsync(dst, b->readback.ptr, bytes, hipMemcpyDeviceToHost, b->stream));
    HB_BTRY(hipStreamSynchronize(b->stream));
    return HB_OK;
}

int hb_batch_read_registers(hb_batch *b, int graph, int which, uint8_t *dst)
{
    if (!b || !dst || graph < 0 || graph >= (int)b->hg.size() || (which != 0 && which != 1))
        return bfail(HB_EINVAL, "hb_batch_read_registers: bad arguments");
    if (!b->ran) return bfail(HB_ESTATE, "hb_batch_read_registers: no completed run");
    HB_BTRY(hipSetDevice(b->device));
    return batch_copy_out(b, graph, batch_regs(b, graph, which), dst, (size_t)1 << b->hg[graph].p);
}

int hb_batch_read_estimates(hb_batch *b, int graph, int which, float *dst)
{
    if (!b || !dst || graph < 0 || graph >= (int)b->hg.size() || (which != 0 && which != 1))
        return bfail(HB_EINVAL, "hb_batch_read_estimates: bad arguments");
    if (!b->ran) return bfail(HB_ESTATE, "hb_batch_read_estimates: no completed run");
    HB_BTRY(hipSetDevice(b->device));
    HB_BTRY(hb::launch_estimates(graph_dev(b, graph), batch_regs(b, graph, which), b->est_tmp.ptr, b->stream));
    return batch_copy_out(b, graph, b->est_tmp.ptr, dst, 4);
}

int hb_batch_read_accumulators(hb_batch *b, int graph, double *dist, double *inv)
{
    if (!b || graph < 0 || graph >= (int)b->hg.size()) return bfail(HB_EINVAL, "hb_batch_read_accumulators: bad arguments");
    if (!b->ran) return bfail(HB_ESTATE, "hb_batch_read_accumulators: no completed run");
    HB_BTRY(hipSetDevice(b->device));
    const auto &g = b->hg[graph];
    if (dist) {
        const int rc = batch_copy_out(b, graph, b->dist.ptr + g.node_base, dist, 8);
        if (rc != HB_OK) return rc;
    }
    if (inv) {
        const int rc = batch_copy_out(b, graph, b->inv.ptr + g.node_base, inv, 8);
        if (rc != HB_OK) return rc;
    }
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
        b->small_max = s;
        b->mid_max = m;
        b->tile_arcs = t;
        b->sched_ok = false;
        *best = 1e300;
        for (int r = 0; r < 2; ++r) {
            const int rc = hb_batch_run(b, HB_FLAG_GRAPH, &st);
            if (rc != HB_OK) return rc;
            *best = std::min(*best, st.ms_total);
        }
        return HB_OK;
    };
    // 0 = the per-precision defaults of the single-graph engine
    uint32_t knob[3] = {b->small_max, b->mid_max, b->tile_arcs};
    const uint32_t cand[3][4] = {{0, 16, 64, 256}, {0, 256, 1024, 4096}, {256, 512, 1024, 4096}};
    double best = 0.0;
    int rc = trial(knob[0], knob[1], knob[2], &best);
    if (rc != HB_OK) return rc;
    for (int k = 0; k < 3; ++k) {
        for (uint32_t v : cand[k]) {
            if (v == knob[k]) continue;
            uint32_t t[3] = {knob[0], knob[1], knob[2]};
            t[k] = v;
            double ms = 0.0;
            rc = trial(t[0], t[1], t[2], &ms);
            if (rc != HB_OK) return rc;
            if (ms < best) {
                best = ms;
                knob[k] = v;
            }
        }
    }
    b->small_max = knob[0];
    b->mid_max = knob[1];
    b->tile_arcs = knob[2];
    b->sched_ok = false;
    if (small_max_deg) *small_max_deg = knob[0];
    if (mid_max_deg) *mid_max_deg = knob[1];
    if (tile_arcs) *tile_arcs = knob[2];
    if (trial_ms) *trial_ms = best;
    b->ran = false;
    hb::set_error("");
    return HB_OK;
}
