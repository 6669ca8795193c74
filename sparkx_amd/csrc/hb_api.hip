// hb_api.hip -- C ABI of the MI355X HyperBall engine (declared in include/sparkx_hb.h).
//
// A context owns one HIP stream and the device-resident state of one graph:
//   CSR (u32 offsets[n+1], u32 neighbours[arcs]) staged once through pinned
//   host buffers; per run: two u8 register buffers [n][2^p], f32 estimates,
//   f64 distance / inverse-distance accumulators, two changed-node bitmaps,
//   the iteration state and per-iteration counters, and the degree-binned
//   schedule (node lists per team width, hub chunk tasks and their scratch).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sparkx_hb.h"
#include "hb_device.h"
#include "hb_kernels.h"

#define HB_BIAS_TABLE_QUAL static const
#include "hll_bias_table.inc"

static_assert(HB_BIAS_K - 5 == HB_BIAS_WINDOWS, "bias windows of 6 consecutive points");
static_assert(sizeof(HbBiasTable) % 8 == 0, "HbBiasTable is staged as 8-byte words");

namespace hb {
namespace {
// Start of the window of 6 consecutive bias points HLL++ EstimateBias averages for the
// estimate e: the first point above e, then 6 steps each taking the nearer of the
// points on either side of the window (the lower one on a tie), in f64 as the rule
// compares them.
int bias_window_start(const double *raw, double e)
{
    int lo = 0, hi = HB_BIAS_K;
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (raw[mid] > e) hi = mid;
        else lo = mid + 1;
    }
    hi = lo;
    for (int k = 0; k < 6; ++k) {
        if (lo > 0 && (hi == HB_BIAS_K || e - raw[lo - 1] <= raw[hi] - e)) --lo;
        else ++hi;
    }
    return lo;
}

void build_bias_table(int p, HbBiasTable *t)
{
    const double *raw = HB_BIAS_RAW[p - HB_BIAS_PMIN], *val = HB_BIAS_VAL[p - HB_BIAS_PMIN];
    memset(t, 0, sizeof *t);
    for (int s = 0; s < HB_BIAS_WINDOWS; ++s) {  // summed in ascending order, as the rule does
        double sum = 0.0;
        for (int i = s; i < s + 6; ++i) sum += val[i];
        t->mean[s] = sum / 6.0;
    }
    // bp[s]: smallest f32 e >= 0 whose window starts at s or later (bisection over the
    // f32 bit patterns, which order the non-negative floats)
    for (int s = 1; s < HB_BIAS_WINDOWS; ++s) {
        uint32_t lo = 0, hi = 0x7f7fffffu;  // start(0) < s <= start(FLT_MAX) = 194
        while (lo + 1 < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            float f;
            memcpy(&f, &mid, 4);
            if (bias_window_start(raw, (double)f) >= s) hi = mid;
            else lo = mid;
        }
        float f;
        memcpy(&f, &hi, 4);
        t->bp[s] = f;
    }
    t->bucket_base = (uint32_t)(127 + p - 1) << 8;
    for (uint32_t b = 0; b < HB_BIAS_BUCKETS; ++b) {
        const uint32_t bits = (t->bucket_base + b) << 15;  // lower edge of bucket b
        float f;
        memcpy(&f, &bits, 4);
        t->bucket[b] = (uint8_t)bias_window_start(raw, (double)f);
    }
}
}  // namespace

// per-precision tables, built once per process
const HbBiasTable *bias_table(int p)
{
    static std::mutex mu;
    static std::unique_ptr<HbBiasTable> cache[HB_BIAS_PMAX + 1];
    if (p < HB_BIAS_PMIN || p > HB_BIAS_PMAX) return nullptr;
    std::lock_guard<std::mutex> lk(mu);
    if (!cache[p]) {
        auto t = std::make_unique<HbBiasTable>();
        build_bias_table(p, t.get());
        cache[p] = std::move(t);
    }
    return cache[p].get();
}
}  // namespace hb

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HB_TRY(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail(e_ == hipErrorOutOfMemory ? HB_ENOMEM : HB_EHIP, "%s failed: %s", #call, \
                        hipGetErrorString(e_));                                               \
    } while (0)

// HLL++ linear-counting thresholds (Heule, Nunkesser, Hall 2013), p = 4..16
const double kLcThreshold[17] = {0, 0, 0, 0, 10, 20, 40, 80, 220, 400, 900,
                                 1800, 3100, 6500, 11500, 20000, 50000};

template <class T>
struct DevBuf {
    T *ptr = nullptr;
    size_t count = 0;
    hipError_t ensure(size_t n)
    {
        if (n <= count && ptr) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&ptr, std::max<size_t>(n, 1) * sizeof(T));
        if (e == hipSuccess) count = n;
        else ptr = nullptr;
        return e;
    }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        count = 0;
    }
    void adopt(T *p, size_t n)  // takes a hipMalloc'd buffer
    {
        release();
        ptr = p;
        count = n;
    }
};

}  // namespace

void hb::set_error(const char *msg) { g_err = msg; }

struct hb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t n = 0, arcs = 0;
    DevBuf<uint32_t> off, nbr;
    std::vector<uint32_t> h_off;  // host copy of the offsets, for scheduling

    // schedule knobs
    // degree bounds of the bins; until hb_set_schedule is called they follow p
    // (default_schedule)
    uint32_t small_max = 256, mid_max = 2048, wave_max = 4096, chunk = 8192;
    bool sched_user = false;

    // run state
    int p = 0;
    bool ran = false;
    DevBuf<uint8_t> regs0, regs1, scratch;
    DevBuf<float> est, est_tmp, lc;
    DevBuf<double> dist, inv, cent;
    DevBuf<HbBiasTable> bias;
    DevBuf<uint32_t> bits0, bits1, sched_buf, chunk_flag;
    DevBuf<uint32_t> active, dyn;  // push iterations: bitmap; per-bin lists + 3 counts
    bool sym = false;              // staged CSR is symmetric (arc multiset fingerprints)
    DevBuf<HbIterState> st;
    DevBuf<HbIterRec> rec;
    hb::HbSchedule sched;
    HbDev dev{};
    HbIterState h_st{};
    std::vector<HbIterRec> h_rec;
    uint32_t sched_p = 0, sched_key[4] = {};
    uint32_t n_linked = 0;
    uint32_t snap_k = 0;      // snapshots the next run records (hb_set_snapshots)
    uint32_t run_snap_k = 0;  // snapshots the completed run recorded (rows in snaps)
    DevBuf<float> snaps, feat, pred;
    DevBuf<uint64_t> walk;      // walk-count passes: S_1, S_2, S_3, M_1, hub partials
    DevBuf<uint32_t> walk_hub;  // hub chunk table of the walk passes, then the u32 degrees
    uint32_t walk_chunks = 0, walk_hubs = 0;  // of the table in walk_hub (valid while allocated)

    // node order on the device (HB_LAYOUT_*), built when the graph is staged; reads
    // return the caller's order.  `layout` is what was asked for, `layout_used` what
    // the staged graph got (HB_LAYOUT_AUTO resolves to CHAIN or DEGREE)
    int layout = HB_LAYOUT_AUTO;
    int layout_used = HB_LAYOUT_INPUT;
    uint32_t chain_levels = 0;  // breadth-first levels the chain layout searched
    bool laid_out = false;  // perm .. rorig hold that numbering of the staged graph
    DevBuf<uint32_t> perm, sigma, roff, rnbr, rorig;
    std::vector<uint32_t> h_roff;
    float ms_layout = 0.0f;  // device time of building it
    DevBuf<uint8_t> readback;  // per-node results moved back to the caller's order

    bool relabelled() const { return layout_used != HB_LAYOUT_INPUT && laid_out; }
    const uint32_t *run_off() const { return relabelled() ? roff.ptr : off.ptr; }
    const uint32_t *run_nbr() const { return relabelled() ? rnbr.ptr : nbr.ptr; }
    const uint32_t *run_orig() const { return relabelled() ? rorig.ptr : nbr.ptr; }
    const uint32_t *run_node_id() const { return relabelled() ? perm.ptr : nullptr; }
    const std::vector<uint32_t> &run_hoff() const { return relabelled() ? h_roff : h_off; }

    void release_layout()
    {
        perm.release();
        sigma.release();
        roff.release();
        rnbr.release();
        rorig.release();
        h_roff.clear();
        laid_out = false;
        layout_used = HB_LAYOUT_INPUT;
        chain_levels = 0;
        ms_layout = 0.0f;
    }

    void release_run()
    {
        regs0.release();
        regs1.release();
        scratch.release();
        est.release();
        est_tmp.release();
        lc.release();
        dist.release();
        inv.release();
        bias.release();
        cent.release();
        bits0.release();
        bits1.release();
        sched_buf.release();
        chunk_flag.release();
        active.release();
        dyn.release();
        st.release();
        rec.release();
        snaps.release();
        feat.release();
        pred.release();
        walk.release();
        walk_hub.release();
        readback.release();
        ran = false;
        sched_p = 0;
    }
};

static bool set_device(hb_ctx *c) { return hipSetDevice(c->device) == hipSuccess; }

static int ensure_layout(hb_ctx *c);

// The CSR is in c->off / c->nbr: keeps a host copy of the offsets (h_off, or read
// back when null), tests symmetry with the arc-multiset fingerprints (symmetric graphs
// may use push iterations: a changed node marks its neighbours) and builds the
// degree layout.
static int finish_stage(hb_ctx *c, uint64_t n, uint64_t arcs, std::vector<uint32_t> *h_off)
{
    c->sym = true;
    if (n) {
        unsigned long long *acc = nullptr, h_acc[4] = {};
        HB_TRY(hipMalloc(&acc, (4 + 4 * hb::kFingerprintBlocks) * sizeof(unsigned long long)));
        hipError_t e = hb::arc_fingerprints(c->off.ptr, c->nbr.ptr, (uint32_t)n, arcs, acc, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(h_acc, acc, sizeof h_acc, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        (void)hipFree(acc);
        if (e != hipSuccess) return fail(HB_EHIP, "staging: symmetry check failed: %s", hipGetErrorString(e));
        c->sym = h_acc[0] == h_acc[1] && h_acc[2] == h_acc[3];
    }
    std::vector<uint32_t> local;
    if (!h_off) {
        local.resize(n + 1);
        HB_TRY(hipMemcpy(local.data(), c->off.ptr, (n + 1) * 4, hipMemcpyDeviceToHost));
        h_off = &local;
    }
    c->h_off.swap(*h_off);
    c->n = n;
    c->arcs = arcs;
    return ensure_layout(c);  // degree layout: built with the staging
}

extern "C" {

const char *hb_last_error(void) { return g_err.c_str(); }

#ifndef HB_BUILD_DIGEST
#define HB_BUILD_DIGEST "unknown"
#endif
const char *hb_build_digest(void) { return HB_BUILD_DIGEST; }

int hb_create(hb_ctx **out, int device_ordinal)
{
    if (!out) return fail(HB_EINVAL, "hb_create: out is NULL");
    *out = nullptr;
    int count = 0;
    HB_TRY(hipGetDeviceCount(&count));
    if (device_ordinal < 0 || device_ordinal >= count)
        return fail(HB_EINVAL, "hb_create: device %d out of range (%d devices)", device_ordinal, count);
    HB_TRY(hipSetDevice(device_ordinal));
    hb_ctx *c = new hb_ctx();
    c->device = device_ordinal;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(HB_EHIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    *out = c;
    g_err.clear();
    return HB_OK;
}

void hb_destroy(hb_ctx *c)
{
    if (!c) return;
    set_device(c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->release_run();
    c->release_layout();
    c->off.release();
    c->nbr.release();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

uint64_t hb_num_nodes(const hb_ctx *c) { return c ? c->n : 0; }
uint64_t hb_num_arcs(const hb_ctx *c) { return c ? c->arcs : 0; }

int hb_set_schedule(hb_ctx *c, uint32_t small_max_deg, uint32_t wave_max_deg, uint32_t chunk)
{
    if (!c) return fail(HB_EINVAL, "hb_set_schedule: ctx is NULL");
    c->sched_user = true;
    if (small_max_deg) c->small_max = small_max_deg;
    if (wave_max_deg) c->wave_max = wave_max_deg;
    if (chunk) c->chunk = chunk;
    if (c->wave_max < c->small_max) c->wave_max = c->small_max;
    c->mid_max = std::max(c->small_max, std::min<uint32_t>(c->wave_max, 8 * c->small_max));
    c->sched_p = 0;  // rebuild on the next run
    return HB_OK;
}

int hb_set_layout(hb_ctx *c, int layout)
{
    if (!c) return fail(HB_EINVAL, "hb_set_layout: ctx is NULL");
    if (layout != HB_LAYOUT_INPUT && layout != HB_LAYOUT_DEGREE && layout != HB_LAYOUT_CHAIN &&
        layout != HB_LAYOUT_AUTO)
        return fail(HB_EINVAL, "hb_set_layout: bad layout %d", layout);
    if (layout != c->layout) {
        if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
        (void)hipStreamSynchronize(c->stream);
        c->release_run();  // results and schedule belong to the old numbering
        c->layout = layout;
        if (int rc = ensure_layout(c)) return rc;  // a staged graph gets it now
    }
    g_err.clear();
    return HB_OK;
}

}  // extern "C"

namespace {
// Two pinned host buffers shared by every staging call of the process (allocated on
// first use, kept to the end: pinning host pages costs milliseconds per call);
// g_pin_mu serialises the calls that use them.
constexpr size_t kPinElems = size_t(8) << 20;  // u32 per buffer (32 MiB)
std::mutex g_pin_mu;
uint32_t *g_pin[2] = {nullptr, nullptr};

// dst[i] = src[pos + i] as u32 for i < len, split over host threads; returns the
// first arc index whose neighbour is >= n, or UINT64_MAX
uint64_t convert_neighbours(const void *nbrs, int nbr_bytes, uint64_t pos, uint64_t len, uint64_t n,
                            uint32_t *dst)
{
    auto run = [&](uint64_t a, uint64_t b) -> uint64_t {
        uint64_t bad = UINT64_MAX;
        if (nbr_bytes == 8) {
            const uint64_t *src = (const uint64_t *)nbrs + pos;
            for (uint64_t i = a; i < b; ++i) {
                const uint64_t v = src[i];
                if (v >= n && bad == UINT64_MAX) bad = pos + i;
                dst[i] = (uint32_t)v;
            }
        } else {
            const uint32_t *src = (const uint32_t *)nbrs + pos;
            uint32_t hi = 0;
            for (uint64_t i = a; i < b; ++i) {
                const uint32_t v = src[i];
                hi = std::max(hi, v);
                dst[i] = v;
            }
            if (hi >= n)
                for (uint64_t i = a; i < b; ++i)
                    if (src[i] >= n) return pos + i;
        }
        return bad;
    };
    const unsigned hw = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    const unsigned T = len >= (1u << 20) ? hw : 1u;
    if (T == 1) return run(0, len);
    std::vector<uint64_t> bad(T, UINT64_MAX);
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < T; ++t)
        pool.emplace_back([&, t] { bad[t] = run(len * t / T, len * (t + 1) / T); });
    for (auto &th : pool) th.join();
    return *std::min_element(bad.begin(), bad.end());
}
}  // namespace

extern "C" {

// Copies offsets/neighbours into HBM through two pinned staging buffers, converting
// to u32 and validating on the way (every neighbour id < n, offsets monotone).
int hb_stage_csr(hb_ctx *c, const void *offsets, int off_bytes, uint64_t n, const void *nbrs,
                 int nbr_bytes, uint64_t arcs)
{
    if (!c) return fail(HB_EINVAL, "hb_stage_csr: ctx is NULL");
    if (!offsets || (arcs && !nbrs)) return fail(HB_EINVAL, "hb_stage_csr: NULL buffer");
    if ((off_bytes != 4 && off_bytes != 8) || (nbr_bytes != 4 && nbr_bytes != 8))
        return fail(HB_EINVAL, "hb_stage_csr: element sizes must be 4 or 8 bytes");
    if (n >= (1ull << 32) - 1 || arcs >= (1ull << 32))
        return fail(HB_ERANGE, "hb_stage_csr: n=%llu arcs=%llu exceed the u32 device layout",
                    (unsigned long long)n, (unsigned long long)arcs);
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    auto off_at = [&](uint64_t i) -> uint64_t {
        return off_bytes == 8 ? ((const uint64_t *)offsets)[i] : ((const uint32_t *)offsets)[i];
    };
    if (off_at(0) != 0 || off_at(n) != arcs)
        return fail(HB_EINVAL, "hb_stage_csr: offsets[0]=%llu offsets[n]=%llu, expected 0 and %llu",
                    (unsigned long long)off_at(0), (unsigned long long)off_at(n),
                    (unsigned long long)arcs);
    std::vector<uint32_t> h_off(n + 1);
    for (uint64_t i = 0; i <= n; ++i) {
        const uint64_t o = off_at(i);
        if (i && o < h_off[i - 1])
            return fail(HB_EINVAL, "hb_stage_csr: offsets decrease at node %llu", (unsigned long long)i - 1);
        h_off[i] = (uint32_t)o;
    }
    c->release_run();
    c->release_layout();
    c->n = 0;
    c->arcs = 0;
    HB_TRY(c->off.ensure(n + 1));
    HB_TRY(c->nbr.ensure(arcs));
    HB_TRY(hipMemcpyAsync(c->off.ptr, h_off.data(), (n + 1) * 4, hipMemcpyHostToDevice, c->stream));

    int rc = HB_OK;
    if (arcs) {
        std::lock_guard<std::mutex> lock(g_pin_mu);
        hipEvent_t done[2] = {nullptr, nullptr};
        for (int b = 0; b < 2 && rc == HB_OK; ++b) {
            if (!g_pin[b] && hipHostMalloc((void **)&g_pin[b], kPinElems * 4, hipHostMallocPortable) != hipSuccess) {
                g_pin[b] = nullptr;
                rc = fail(HB_ENOMEM, "hipHostMalloc of %zu bytes failed", kPinElems * 4);
            } else if (hipEventCreateWithFlags(&done[b], hipEventDisableTiming) != hipSuccess) {
                rc = fail(HB_EHIP, "hipEventCreate failed");
            }
        }
        uint64_t pos = 0;
        int b = 0;
        while (rc == HB_OK && pos < arcs) {
            const uint64_t len = std::min<uint64_t>(kPinElems, arcs - pos);
            if (hipEventSynchronize(done[b]) != hipSuccess) {  // the copy out of g_pin[b] is done
                rc = fail(HB_EHIP, "hipEventSynchronize failed");
                break;
            }
            const uint64_t bad_at = convert_neighbours(nbrs, nbr_bytes, pos, len, n, g_pin[b]);
            if (bad_at != UINT64_MAX) {
                const uint64_t v = nbr_bytes == 8 ? ((const uint64_t *)nbrs)[bad_at] : ((const uint32_t *)nbrs)[bad_at];
                rc = fail(HB_EINVAL, "hb_stage_csr: neighbour %llu at arc %llu is not < n=%llu", (unsigned long long)v,
                          (unsigned long long)bad_at, (unsigned long long)n);
                break;
            }
            if (hipMemcpyAsync(c->nbr.ptr + pos, g_pin[b], len * 4, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
                hipEventRecord(done[b], c->stream) != hipSuccess) {
                rc = fail(HB_EHIP, "staging copy failed");
                break;
            }
            pos += len;
            b ^= 1;
        }
        (void)hipStreamSynchronize(c->stream);  // no copy still reads the shared buffers
        for (int i = 0; i < 2; ++i)
            if (done[i]) (void)hipEventDestroy(done[i]);
    }
    if (rc != HB_OK) return rc;
    if (int rc2 = finish_stage(c, n, arcs, &h_off)) return rc2;
    g_err.clear();
    return HB_OK;
}

}  // extern "C"

// Default bin bounds, measured with the degree layout on R-MAT-20 / R-MAT-24 (sweeps of
// small_max 16..512, wave_max 256..16384, chunk 2048..8192): at p >= 6 a node of degree
// <= 256 is cheapest on one sketch copy (2^p/16 lanes), at p <= 5 (one or two lanes per
// copy) the first iteration's hashing wants the wider teams from degree 64 on.  Hub
// chunks shrink with the graph so that smaller graphs still spread their hubs over the
// whole GPU: chunk = the power of two >= arcs / 65536 within [2048, 8192] (R-MAT-24:
// 8192, R-MAT-20: 2048), wave_max = max(2048, chunk / 2).  mid_max = min(wave_max,
// 8 small_max).
static void default_schedule(hb_ctx *c, int p)
{
    if (c->sched_user) return;
    c->small_max = p <= 5 ? 64 : 256;
    uint32_t chunk = 2048;
    while (chunk < 8192 && (uint64_t)chunk * 65536 < c->arcs) chunk *= 2;
    c->chunk = chunk;
    c->wave_max = p <= 5 ? 16384 : std::max<uint32_t>(2048, c->chunk / 2);
    c->mid_max = std::max(c->small_max, std::min<uint32_t>(c->wave_max, 8 * c->small_max));
}

// Degree-binned schedule: bins of team width G (deg <= small_max), 4G (<= mid_max)
// and 64 (<= wave_max); higher-degree nodes are split into chunk-sized wavefront
// tasks merged per hub.
static int build_schedule(hb_ctx *c, int p)
{
    const uint32_t key[4] = {c->small_max, c->mid_max, c->wave_max, c->chunk};
    if (c->sched_p == (uint32_t)p && !memcmp(key, c->sched_key, sizeof key)) return HB_OK;
    const uint32_t G = std::min(64u, (1u << p) / 16u);
    const uint32_t n = (uint32_t)c->n;
    const std::vector<uint32_t> &h_off = c->run_hoff();
    c->n_linked = 0;
    std::vector<uint32_t> bins[3], tnode, tbeg, tend, hnode, htask0, hntask;
    for (uint32_t u = 0; u < n; ++u) {
        const uint32_t d = h_off[u + 1] - h_off[u];
        if (d == 0) continue;  // keeps c_0; written to both buffers at init
        ++c->n_linked;
        if (p > 10) bins[2].push_back(u);  // one wavefront per node walks the register row
        else if (d <= c->small_max) bins[0].push_back(u);
        else if (d <= c->mid_max) bins[1].push_back(u);
        else if (d <= c->wave_max) bins[2].push_back(u);
        else {
            hnode.push_back(u);
            htask0.push_back((uint32_t)tnode.size());
            uint32_t cnt = 0;
            for (uint32_t e = h_off[u]; e < h_off[u + 1]; e += c->chunk, ++cnt) {
                tnode.push_back(u);
                tbeg.push_back(e);
                tend.push_back(std::min(e + c->chunk, h_off[u + 1]));
            }
            hntask.push_back(cnt);
        }
    }
    const size_t total = bins[0].size() + bins[1].size() + bins[2].size() + 3 * tnode.size() +
                         3 * hnode.size();
    HB_TRY(c->sched_buf.ensure(total));
    std::vector<uint32_t> host;
    host.reserve(total);
    hb::HbSchedule s;
    size_t at = 0;
    const int teams[3] = {(int)G, (int)std::min(64u, 4 * G), 64};
    for (int b = 0; b < 3; ++b) {
        s.bin_count[s.nbins] = (uint32_t)bins[b].size();
        s.bin_team[s.nbins] = teams[b];
        // ids ascend within a bin, so it is one id range exactly when its ends are
        // size - 1 apart (every bin of the degree and chain layouts): the kernels then
        // compute the ids instead of loading them
        if (p <= 10 && !bins[b].empty() && bins[b].back() - bins[b].front() + 1 == bins[b].size()) {
            s.bin_list[s.nbins] = nullptr;
            s.bin_base[s.nbins] = bins[b].front();
        } else {
            s.bin_list[s.nbins] = c->sched_buf.ptr + at;
            host.insert(host.end(), bins[b].begin(), bins[b].end());
            at += bins[b].size();
        }
        ++s.nbins;
    }
    auto put = [&](const std::vector<uint32_t> &v) -> const uint32_t * {
        const uint32_t *ptr = c->sched_buf.ptr + at;
        host.insert(host.end(), v.begin(), v.end());
        at += v.size();
        return ptr;
    };
    s.ntasks = (uint32_t)tnode.size();
    s.task_node = put(tnode);
    s.task_beg = put(tbeg);
    s.task_end = put(tend);
    s.nhubs = (uint32_t)hnode.size();
    s.hub_node = put(hnode);
    s.hub_task0 = put(htask0);
    s.hub_ntask = put(hntask);
    if (!host.empty())
        HB_TRY(hipMemcpyAsync(c->sched_buf.ptr, host.data(), host.size() * 4, hipMemcpyHostToDevice,
                              c->stream));
    HB_TRY(c->scratch.ensure((size_t)s.ntasks << p));
    HB_TRY(c->chunk_flag.ensure(s.ntasks));
    if (c->sym && p <= 10) HB_TRY(c->dyn.ensure((size_t)s.bin_count[0] + s.bin_count[1] + s.bin_count[2] + 3));
    HB_TRY(hipStreamSynchronize(c->stream));
    c->sched = s;
    c->sched_p = (uint32_t)p;
    memcpy(c->sched_key, key, sizeof key);
    return HB_OK;
}

static float alpha_f32(uint32_t m)
{
    if (m == 16) return 0.673f;
    if (m == 32) return 0.697f;
    if (m == 64) return 0.709f;
    return 0.7213f / (1.0f + 1.079f / (float)m);
}

// HB_LAYOUT_AUTO: the chain layout for graphs made of long low-degree paths (every
// degree <= 64 and at most 4 arcs per node on average, as de Bruijn graphs of
// genomes), where it puts a node's neighbours in its own cache lines; the degree
// layout otherwise, where the gathers go to the high-degree nodes' rows.
static int resolve_layout(const hb_ctx *c)
{
    if (c->layout != HB_LAYOUT_AUTO) return c->layout;
    uint32_t max_deg = 0;
    for (size_t u = 0; u + 1 < c->h_off.size(); ++u) max_deg = std::max(max_deg, c->h_off[u + 1] - c->h_off[u]);
    return max_deg <= 64 && c->arcs <= 4 * (uint64_t)c->n ? HB_LAYOUT_CHAIN : HB_LAYOUT_DEGREE;
}

// Builds the device numbering of the staged graph once (device work, timed with
// events into ms_layout); a no-op for HB_LAYOUT_INPUT or when already built.
static int ensure_layout(hb_ctx *c)
{
    if (c->layout == HB_LAYOUT_INPUT || c->laid_out || c->n == 0) return HB_OK;
    const int kind = resolve_layout(c);
    const uint32_t n = (uint32_t)c->n, arcs = (uint32_t)c->arcs;
    HB_TRY(c->perm.ensure(n));
    HB_TRY(c->sigma.ensure(n));
    HB_TRY(c->roff.ensure((size_t)n + 1));
    HB_TRY(c->rnbr.ensure(arcs));
    HB_TRY(c->rorig.ensure(arcs));
    hipEvent_t e0, e1;
    HB_TRY(hipEventCreate(&e0));
    HB_TRY(hipEventCreate(&e1));
    (void)hipEventRecord(e0, c->stream);
    uint32_t levels = 0;
    hipError_t e = kind == HB_LAYOUT_CHAIN
                       ? hb::relabel_by_chain(c->off.ptr, c->nbr.ptr, n, arcs, c->perm.ptr, c->sigma.ptr, c->roff.ptr,
                                              c->rnbr.ptr, c->rorig.ptr, 4096, &levels, c->stream)
                       : hb::relabel_by_degree(c->off.ptr, c->nbr.ptr, n, arcs, c->perm.ptr, c->sigma.ptr,
                                               c->roff.ptr, c->rnbr.ptr, c->rorig.ptr, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    std::vector<uint32_t> h((size_t)n + 1);
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), c->roff.ptr, ((size_t)n + 1) * 4, hipMemcpyDeviceToHost,
                                            c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0.0f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) {
        c->release_layout();
        return fail(e == hipErrorOutOfMemory ? HB_ENOMEM : HB_EHIP, "%s layout failed: %s",
                    kind == HB_LAYOUT_CHAIN ? "chain" : "degree", hipGetErrorString(e));
    }
    c->h_roff.swap(h);
    c->ms_layout = ms;
    c->layout_used = kind;
    c->chain_levels = levels;
    c->laid_out = true;
    return HB_OK;
}

// Copies n per-node rows of row_bytes from device src to host dst in the caller's
// node order (through c->readback when the nodes are renumbered).  Synchronous.
static int copy_out(hb_ctx *c, const void *src, void *dst, size_t row_bytes)
{
    const size_t bytes = (size_t)c->n * row_bytes;
    if (!c->relabelled()) {
        HB_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
        HB_TRY(hipStreamSynchronize(c->stream));
        return HB_OK;
    }
    HB_TRY(c->readback.ensure(bytes));
    HB_TRY(hb::unpermute_rows(src, c->readback.ptr, c->sigma.ptr, (uint32_t)c->n, row_bytes, c->stream));
    HB_TRY(hipMemcpyAsync(dst, c->readback.ptr, bytes, hipMemcpyDeviceToHost, c->stream));
    HB_TRY(hipStreamSynchronize(c->stream));
    return HB_OK;
}

extern "C" int hb_run(hb_ctx *c, int p, int max_iters, uint32_t flags, hb_stats *stats)
{
    if (!c) return fail(HB_EINVAL, "hb_run: ctx is NULL");
    if (p < 4 || p > 16) return fail(HB_EINVAL, "hb_run: precision p=%d outside 4..16", p);
    if (max_iters < 0) return fail(HB_EINVAL, "hb_run: max_iters < 0");
    if (!c->off.ptr) return fail(HB_ESTATE, "hb_run: no graph staged");
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    const uint32_t n = (uint32_t)c->n;
    const uint32_t m = 1u << p;
    const uint64_t words = (c->n + 31) / 32;
    c->ran = false;
    if (n == 0) {
        // no nodes: the reference's loop runs one pass with no increase and stops
        c->h_st = HbIterState{};
        c->h_st.executed = 1;
        c->h_st.converged = 1;
        c->h_rec.assign(2, HbIterRec{});
        c->p = p;
        c->run_snap_k = 0;
        c->ran = true;
        if (stats) {
            memset(stats, 0, sizeof *stats);
            stats->iterations = 1;
            stats->converged = 1;
            stats->p = p;
        }
        g_err.clear();
        return HB_OK;
    }

    HB_TRY(c->regs0.ensure((size_t)c->n << p));
    HB_TRY(c->regs1.ensure((size_t)c->n << p));
    HB_TRY(c->est.ensure(c->n));
    HB_TRY(c->dist.ensure(c->n));
    HB_TRY(c->inv.ensure(c->n));
    HB_TRY(c->bits0.ensure(words));
    HB_TRY(c->bits1.ensure(words));
    HB_TRY(c->st.ensure(1));
    // No iteration cap: a run ends by the reference's rule or at max_iters.  A graph
    // cannot change a register after its diameter, so n + 2 iterations bound a correct
    // run; the host loop stops there with an error rather than going on.
    const uint64_t bound = max_iters > 0 ? (uint64_t)max_iters : (uint64_t)c->n + 2;
    HB_TRY(c->rec.ensure(HB_REC_RING));
    HB_TRY(c->lc.ensure(m + 1));
    HB_TRY(c->bias.ensure(1));
    if (c->snap_k) HB_TRY(c->snaps.ensure((size_t)c->snap_k * c->n));
    const bool push_ok = c->sym && p <= 10;
    if (push_ok) HB_TRY(c->active.ensure(2 * (size_t)words));  // two, by iteration parity
    {
        int rc = ensure_layout(c);  // normally built at staging already
        default_schedule(c, p);
        if (rc == HB_OK) rc = build_schedule(c, p);
        if (rc != HB_OK) return rc;
    }

    // host-computed constants: LC table (f64 then f32, as m ln(m/z)), alpha m^2, bias points
    std::vector<float> lc(m + 1, 0.0f);
    for (uint32_t z = 1; z <= m; ++z) lc[z] = (float)((double)m * std::log((double)m / (double)z));
    HB_TRY(hipMemcpyAsync(c->lc.ptr, lc.data(), (m + 1) * 4, hipMemcpyHostToDevice, c->stream));
    HB_TRY(hipMemcpyAsync(c->bias.ptr, hb::bias_table(p), sizeof(HbBiasTable), hipMemcpyHostToDevice,
                          c->stream));

    HbDev &d = c->dev;
    d = HbDev{};
    d.off = c->run_off();
    d.nbr = c->run_nbr();
    d.orig = c->run_orig();
    d.node_id = c->run_node_id();
    d.regs0 = c->regs0.ptr;
    d.regs1 = c->regs1.ptr;
    d.est = c->est.ptr;
    d.dist = c->dist.ptr;
    d.inv = c->inv.ptr;
    d.bits0 = c->bits0.ptr;
    d.bits1 = c->bits1.ptr;
    d.st = c->st.ptr;
    d.rec = c->rec.ptr;
    d.lc = c->lc.ptr;
    d.bias = c->bias.ptr;
    d.scratch = c->scratch.ptr;
    d.chunk_flag = c->chunk_flag.ptr;
    d.sym = push_ok ? 1 : 0;
    d.arcs = c->arcs;
    d.small_max = c->small_max;
    d.mid_max = c->mid_max;
    d.wave_max = c->wave_max;
    if (push_ok) {
        d.active = c->active.ptr;
        d.dyn_list = c->dyn.ptr + 3;
        d.dyn_count = c->dyn.ptr;
        d.dyn_base[0] = 0;
        d.dyn_base[1] = c->sched.bin_count[0];
        d.dyn_base[2] = c->sched.bin_count[0] + c->sched.bin_count[1];
    }
    d.snaps = c->snap_k ? c->snaps.ptr : nullptr;
    d.snap_k = c->snap_k;
    d.n = n;
    d.n_linked = c->n_linked;
    d.p = (uint32_t)p;
    const float mm = (float)m * (float)m;
    d.am = alpha_f32(m) * mm;
    d.five_m = 5.0f * (float)m;
    d.lc_thr = kLcThreshold[p];
    d.frontier = (flags & HB_FLAG_DENSE) ? 0 : 1;
    d.window = c->relabelled() && c->layout_used == HB_LAYOUT_CHAIN ? 1 : 0;
    d.rows = d.window && (p == 5 || p == 6) ? 1 : 0;

    HbIterState st0{};
    st0.iter = 1;
    st0.max_iters = (uint32_t)max_iters;
    st0.t = 1.0;
    st0.inv_t = 1.0;
    HB_TRY(hipMemcpyAsync(c->st.ptr, &st0, sizeof st0, hipMemcpyHostToDevice, c->stream));
    HB_TRY(hb::launch_fill(c->bits0.ptr, words, 0u, c->stream));
    HB_TRY(hb::launch_fill(c->bits1.ptr, words, 0u, c->stream));
    // iteration t marks the active bitmap iteration t - 1 cleared; both start clear
    if (push_ok) HB_TRY(hb::launch_fill(c->active.ptr, 2 * (uint64_t)words, 0u, c->stream));

    hipEvent_t e0, e1, e2;
    HB_TRY(hipEventCreate(&e0));
    HB_TRY(hipEventCreate(&e1));
    HB_TRY(hipEventCreate(&e2));
    std::vector<float> iter_ms;
    int rc = HB_OK;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    const bool nf = (flags & HB_FLAG_NF) != 0;
    const bool ktiming = (flags & HB_FLAG_KTIMING) != 0;
    const bool timing = ktiming || (flags & HB_FLAG_TIMING) != 0;
    if (ktiming && (flags & HB_FLAG_GRAPH)) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        return fail(HB_EINVAL, "hb_run: HB_FLAG_KTIMING cannot be combined with HB_FLAG_GRAPH");
    }
    // timing: each launch of a check's batch has its own pair of events around the
    // iteration and, with HB_FLAG_KTIMING, its own events around each kernel slot; they
    // are read after the check, so timing adds no host round trip per iteration
    constexpr uint32_t kMaxBatch = 8;  // launches between two checks
    const uint32_t per_it = 2 + (ktiming ? 2 * HB_NSLOT : 0);
    std::vector<hipEvent_t> tev(timing ? (size_t)kMaxBatch * per_it : 0, nullptr);
    double kms[HB_NSLOT] = {};
    uint64_t klaunch[HB_NSLOT] = {};
    do {
        bool events_ok = true;
        for (auto &ev : tev) events_ok = events_ok && hipEventCreate(&ev) == hipSuccess;
        if (!events_ok) { rc = fail(HB_EHIP, "event create failed"); break; }
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail(HB_EHIP, "event record failed"); break; }
        hipError_t e = hb::launch_init(d, c->stream);
        if (e != hipSuccess) { rc = fail(HB_EHIP, "init kernel: %s", hipGetErrorString(e)); break; }
        if (hipEventRecord(e1, c->stream) != hipSuccess) { rc = fail(HB_EHIP, "event record failed"); break; }
        if (flags & HB_FLAG_GRAPH) {
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                rc = fail(HB_EHIP, "hipStreamBeginCapture failed");
                break;
            }
            e = hb::launch_iteration(d, c->sched, nf, c->stream);
            hipError_t e2c = hipStreamEndCapture(c->stream, &graph);
            if (e != hipSuccess || e2c != hipSuccess) {
                rc = fail(HB_EHIP, "graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2c));
                break;
            }
            if (hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0) != hipSuccess) {
                rc = fail(HB_EHIP, "hipGraphInstantiate failed");
                break;
            }
        }
        // iterations: launches after convergence return at once (done flag), so a
        // batch can run ahead of the host's check; each check drains the finished
        // iterations' records from the device ring into c->h_rec (index t)
        uint64_t launched = 0;
        HbIterState hs{};
        uint32_t drained = 0;
        HbIterRec ring[HB_REC_RING];
        c->h_rec.assign(1, HbIterRec{});
        for (;;) {
            const uint32_t k = launched < 8 ? 2u : kMaxBatch;
            if (launched >= bound + 16) {
                rc = fail(HB_ESTATE, "hb_run: still changing after %llu iterations", (unsigned long long)launched);
                break;
            }
            for (uint32_t i = 0; i < k && rc == HB_OK; ++i) {
                hipEvent_t *ev = timing ? &tev[(size_t)i * per_it] : nullptr;
                if (ev) (void)hipEventRecord(ev[0], c->stream);
                e = gexec ? hipGraphLaunch(gexec, c->stream)
                          : hb::launch_iteration(d, c->sched, nf, c->stream, ktiming ? ev + 2 : nullptr);
                if (e != hipSuccess) rc = fail(HB_EHIP, "iteration launch: %s", hipGetErrorString(e));
                if (ev) (void)hipEventRecord(ev[1], c->stream);
            }
            launched += k;
            if (rc != HB_OK) break;
            if (hipMemcpyAsync(&hs, c->st.ptr, sizeof hs, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipMemcpyAsync(ring, c->rec.ptr, sizeof ring, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = fail(HB_EHIP, "iteration failed: %s", hipGetErrorString(hipGetLastError()));
                break;
            }
            const uint32_t finished = hs.done ? hs.executed : hs.iter - 1;
            for (uint32_t t = drained + 1; t <= finished; ++t) c->h_rec.push_back(ring[t % HB_REC_RING]);
            drained = finished;
            if (timing) {
                const uint32_t counts[HB_NSLOT] = {c->sched.bin_count[0], c->sched.bin_count[1],
                                                   c->sched.bin_count[2], c->sched.ntasks, c->sched.nhubs};
                for (uint32_t i = 0; i < k; ++i) {
                    // launch number launched - k + i + 1 ran iteration launched - k + i + 1
                    // unless the run had stopped before it (it then returned at once)
                    if (hs.done && launched - k + i + 1 > hs.executed) break;
                    const hipEvent_t *ev = &tev[(size_t)i * per_it];
                    float ms = 0.0f;
                    (void)hipEventElapsedTime(&ms, ev[0], ev[1]);
                    iter_ms.push_back(ms);
                    if (!ktiming) continue;
                    for (int sl = 0; sl < HB_NSLOT; ++sl) {
                        if (!counts[sl]) continue;
                        float kms1 = 0.0f;
                        (void)hipEventElapsedTime(&kms1, ev[2 + 2 * sl], ev[3 + 2 * sl]);
                        kms[sl] += kms1;
                        klaunch[sl] += 1;
                    }
                }
            }
            if (hs.done) break;
        }
        if (rc != HB_OK) break;
        if (hipEventRecord(e2, c->stream) != hipSuccess || hipEventSynchronize(e2) != hipSuccess) {
            rc = fail(HB_EHIP, "event sync failed");
            break;
        }
        c->h_st = hs;
    } while (0);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto &ev : tev)
        if (ev) (void)hipEventDestroy(ev);
    float ms_init = 0.0f, ms_loop = 0.0f;
    if (rc == HB_OK) {
        (void)hipEventElapsedTime(&ms_init, e0, e1);
        (void)hipEventElapsedTime(&ms_loop, e1, e2);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    if (rc != HB_OK) return rc;

    const uint32_t executed = c->h_st.executed;
    // snapshot rows past the last iteration repeat its estimates (the sketches stopped changing)
    for (uint32_t t = executed + 1; t <= c->snap_k; ++t)
        HB_TRY(hipMemcpyAsync(c->snaps.ptr + (size_t)(t - 1) * c->n, c->est.ptr, c->n * 4, hipMemcpyDeviceToDevice,
                              c->stream));
    c->p = p;
    c->run_snap_k = c->snap_k;
    c->ran = true;
    if (stats) {
        memset(stats, 0, sizeof *stats);
        stats->iterations = (int32_t)executed;
        stats->converged = (int32_t)c->h_st.converged;
        // the reference's counters file is buffer 0; the last iteration wrote buffer parity^1
        stats->final_parity_latest = (c->h_st.parity ^ 1u) == 0u;
        stats->p = p;
        stats->ms_total = ms_loop;
        stats->ms_init = ms_init;
        const uint32_t items[HB_NSLOT] = {c->sched.bin_count[0], c->sched.bin_count[1],
                                          c->sched.bin_count[2], c->sched.ntasks, c->sched.nhubs};
        for (int sl = 0; sl < HB_NSLOT; ++sl) {
            stats->kernel_launches[sl] = ktiming ? klaunch[sl] : (items[sl] ? executed : 0);
            stats->kernel_ms[sl] = kms[sl];
            stats->kernel_items[sl] = items[sl];
            stats->kernel_team[sl] = sl < 3 ? (uint32_t)c->sched.bin_team[sl] : 64u;
        }
        if (p <= 10 && executed >= 1)
            for (int sl = 0; sl < HB_NSLOT; ++sl) {
                // bins hash the ids of the neighbours and of the node itself, hub chunks
                // those of the neighbours; the hub merge reads register blocks as always,
                // and so does a bin that k_union_rows runs (it gathers the c_0 rows)
                if (sl < HB_SLOT_CHUNK && hb::rows_bin(c->dev, c->sched, sl)) continue;
                if (sl <= HB_SLOT_CHUNK) stats->kernel_first_arcs[sl] = c->h_rec[1].arcs[sl];
                if (sl < HB_SLOT_CHUNK) stats->kernel_first_nodes[sl] = c->h_rec[1].nodes[sl];
            }
        stats->symmetric = c->sym ? 1 : 0;
        stats->layout = c->relabelled() ? c->layout_used : HB_LAYOUT_INPUT;
        stats->ms_layout = c->relabelled() ? c->ms_layout : 0.0f;
        for (uint32_t t = 1; t <= executed; ++t) {
            const HbIterRec &r = c->h_rec[t];
            stats->push_iterations += r.push ? 1 : 0;
            uint64_t inc = 0, arcs = 0, nodes = 0;
            for (int sl = 0; sl < HB_NSLOT; ++sl) {
                inc += r.inc[sl];
                arcs += r.arcs[sl];
                nodes += r.nodes[sl];
                stats->kernel_arcs[sl] += r.arcs[sl];
                stats->kernel_nodes[sl] += r.nodes[sl];
                stats->kernel_gathered[sl] += r.gathered[sl];
            }
            stats->arcs_gathered += arcs;
            stats->nodes_recomputed += nodes;
            if (t <= HB_STATS_MAX_ITERS) {
                stats->increases[t - 1] = inc;
                stats->iter_arcs[t - 1] = arcs;
                stats->iter_nodes[t - 1] = nodes;
                if (t - 1 < iter_ms.size()) stats->iter_ms[t - 1] = iter_ms[t - 1];
                if (nf) {
                    const unsigned __int128 v = ((unsigned __int128)r.nf_hi << 64) | r.nf_lo;
                    stats->nf[t - 1] = std::ldexp((double)v, -40);
                }
            }
        }
    }
    g_err.clear();
    return HB_OK;
}

static const uint8_t *select_regs(hb_ctx *c, int which)
{
    if (which == HB_REGS_REFERENCE) return c->regs0.ptr;
    return (c->h_st.parity ^ 1u) ? c->regs1.ptr : c->regs0.ptr;
}

extern "C" int hb_read_registers(hb_ctx *c, int which, uint8_t *dst)
{
    if (!c || !dst) return fail(HB_EINVAL, "hb_read_registers: NULL argument");
    if (!c->ran) return fail(HB_ESTATE, "hb_read_registers: no completed run");
    if (which != HB_REGS_REFERENCE && which != HB_REGS_LATEST)
        return fail(HB_EINVAL, "hb_read_registers: bad selector %d", which);
    if (c->n == 0) return HB_OK;
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    if (int rc = copy_out(c, select_regs(c, which), dst, (size_t)1 << c->p)) return rc;
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_read_estimates(hb_ctx *c, int which, float *dst)
{
    if (!c || !dst) return fail(HB_EINVAL, "hb_read_estimates: NULL argument");
    if (!c->ran) return fail(HB_ESTATE, "hb_read_estimates: no completed run");
    if (which != HB_REGS_REFERENCE && which != HB_REGS_LATEST)
        return fail(HB_EINVAL, "hb_read_estimates: bad selector %d", which);
    if (c->n == 0) return HB_OK;
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    HB_TRY(c->est_tmp.ensure(c->n));
    HB_TRY(hb::launch_estimates(c->dev, select_regs(c, which), c->est_tmp.ptr, c->stream));
    if (int rc = copy_out(c, c->est_tmp.ptr, dst, 4)) return rc;
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_read_accumulators(hb_ctx *c, double *dist, double *inv)
{
    if (!c) return fail(HB_EINVAL, "hb_read_accumulators: ctx is NULL");
    if (!c->ran) return fail(HB_ESTATE, "hb_read_accumulators: no completed run");
    if (c->n == 0) return HB_OK;
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    if (dist)
        if (int rc = copy_out(c, c->dist.ptr, dist, 8)) return rc;
    if (inv)
        if (int rc = copy_out(c, c->inv.ptr, inv, 8)) return rc;
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_centrality(hb_ctx *c, int kind, int normalize, double *dst)
{
    if (!c || !dst) return fail(HB_EINVAL, "hb_centrality: NULL argument");
    if (kind < HB_CLOSENESS || kind > HB_LIN) return fail(HB_EINVAL, "hb_centrality: bad kind %d", kind);
    if (!c->ran) return fail(HB_ESTATE, "hb_centrality: no completed run");
    if (c->n == 0) return HB_OK;
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    HB_TRY(c->est_tmp.ensure(c->n));
    HB_TRY(c->cent.ensure(c->n));
    // the epilogues re-estimate the reference's counters buffer (hyper_ball.rs:452,506,544)
    HB_TRY(hb::launch_estimates(c->dev, c->regs0.ptr, c->est_tmp.ptr, c->stream));
    HB_TRY(hb::launch_centrality(c->dev, c->est_tmp.ptr, kind, normalize ? 1 : 0, c->cent.ptr, c->stream));
    if (int rc = copy_out(c, c->cent.ptr, dst, 8)) return rc;
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_exact_bfs(hb_ctx *c, uint64_t src_begin, uint64_t src_end, uint64_t *reach, uint64_t *total,
                            double *inv_sum)
{
    if (!c) return fail(HB_EINVAL, "hb_exact_bfs: ctx is NULL");
    if (src_end > src_begin && !reach && !total && !inv_sum) return fail(HB_EINVAL, "hb_exact_bfs: no output");
    if (!c->off.ptr) return fail(HB_ESTATE, "hb_exact_bfs: no graph staged");
    if (src_begin > src_end || src_end > c->n)
        return fail(HB_EINVAL, "hb_exact_bfs: sources [%llu, %llu) outside [0, %llu)", (unsigned long long)src_begin,
                    (unsigned long long)src_end, (unsigned long long)c->n);
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    const int rc = hb::exact_bfs(c->off.ptr, c->h_off.data(), c->nbr.ptr, (uint32_t)c->n, c->sym, src_begin, src_end,
                                 reach, total, inv_sum, c->stream);
    if (rc == HB_OK) g_err.clear();
    return rc;
}

extern "C" int hb_exact_centrality(hb_ctx *c, int kind, int normalization, uint64_t src_begin, uint64_t src_end,
                                   double *dst)
{
    if (!c || (!dst && src_end > src_begin)) return fail(HB_EINVAL, "hb_exact_centrality: NULL argument");
    if (kind < HB_CLOSENESS || kind > HB_LIN) return fail(HB_EINVAL, "hb_exact_centrality: bad kind %d", kind);
    if (normalization < -1 || normalization > 1)
        return fail(HB_EINVAL, "hb_exact_centrality: normalization %d not in -1..1", normalization);
    if (src_end <= src_begin) return src_end < src_begin ? fail(HB_EINVAL, "hb_exact_centrality: empty range") : HB_OK;
    const uint64_t cnt = src_end - src_begin;
    std::vector<uint64_t> reach(cnt), total(cnt);
    std::vector<double> inv(kind == HB_HARMONIC ? cnt : 0);
    const int rc = hb_exact_bfs(c, src_begin, src_end, reach.data(), total.data(),
                                kind == HB_HARMONIC ? inv.data() : nullptr);
    if (rc != HB_OK) return rc;
    // exact.rs:160-201 (closeness), :379-441 (harmonic), :610-619 (Lin); the
    // reference's f64 distance sum equals the integer total below 2^53
    const double norm = c->n ? (double)(c->n - 1) : 0.0;
    for (uint64_t i = 0; i < cnt; ++i) {
        const double r = (double)reach[i], t = (double)total[i];
        double v;
        if (kind == HB_CLOSENESS) {
            const double num = normalization < 0 ? 0.0 : (normalization == 0 ? norm : r);
            v = (reach[i] <= 1 || t == 0.0) ? 0.0 : (normalization < 0 ? t : num / t);
        } else if (kind == HB_HARMONIC) {
            const double num = normalization < 0 ? 0.0 : (normalization == 0 ? norm : r);
            v = reach[i] <= 1 ? 0.0 : (normalization < 0 ? inv[i] : num / inv[i]);
        } else {
            v = t == 0.0 ? 0.0 : (r * r) / t;
        }
        dst[i] = v;
    }
    return HB_OK;
}

extern "C" int hb_neighbourhood_function(hb_ctx *c, double *dst, int cap)
{
    if (!c || (!dst && cap > 0)) return fail(HB_EINVAL, "hb_neighbourhood_function: NULL argument");
    if (!c->ran) return fail(HB_ESTATE, "hb_neighbourhood_function: no completed run");
    const int it = (int)c->h_st.executed;
    const int k = std::min(it, cap);
    for (int t = 1; t <= k; ++t) {
        const HbIterRec &r = c->h_rec[t];
        const unsigned __int128 v = ((unsigned __int128)r.nf_hi << 64) | r.nf_lo;
        dst[t - 1] = std::ldexp((double)v, -40);
    }
    g_err.clear();
    return k;
}

// Coordinate-descent autotuner over the degree-bin and hub-chunk knobs: each
// candidate runs the first `trial_iters` iterations (the dense, most expensive
// ones) twice and keeps its best device time.  Leaves the best schedule set.
extern "C" int hb_autotune(hb_ctx *c, int p, int trial_iters, uint32_t *small_max_deg,
                           uint32_t *wave_max_deg, uint32_t *chunk, double *trial_ms)
{
    if (!c) return fail(HB_EINVAL, "hb_autotune: ctx is NULL");
    if (trial_iters <= 0) trial_iters = 3;
    std::unique_ptr<hb_stats> st(new hb_stats());
    auto trial = [&](uint32_t s, uint32_t w, uint32_t k, double *best) -> int {
        int rc = hb_set_schedule(c, s, w, k);
        if (rc != HB_OK) return rc;
        *best = 1e300;
        for (int r = 0; r < 2; ++r) {
            rc = hb_run(c, p, trial_iters, 0, st.get());
            if (rc != HB_OK) return rc;
            *best = std::min(*best, st->ms_total);
        }
        return HB_OK;
    };
    default_schedule(c, p);  // the search starts from this precision's defaults
    uint32_t knob[3] = {c->small_max, c->wave_max, c->chunk};
    const uint32_t cand[3][4] = {{16, 64, 256, 512}, {1024, 4096, 16384, 65536}, {2048, 4096, 8192, 16384}};
    double best = 0.0;
    int rc = trial(knob[0], knob[1], knob[2], &best);
    if (rc != HB_OK) return rc;
    if (p <= 10) {
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
    }
    rc = hb_set_schedule(c, knob[0], knob[1], knob[2]);
    if (rc != HB_OK) return rc;
    if (small_max_deg) *small_max_deg = knob[0];
    if (wave_max_deg) *wave_max_deg = knob[1];
    if (chunk) *chunk = knob[2];
    if (trial_ms) *trial_ms = best;
    c->ran = false;  // the trial results are not a completed run
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_result_info(const hb_ctx *c, int *p, int *iterations, int *snapshots)
{
    if (!c) return fail(HB_EINVAL, "hb_result_info: ctx is NULL");
    if (!c->ran) return fail(HB_ESTATE, "hb_result_info: no completed run");
    if (p) *p = c->p;
    if (iterations) *iterations = (int)c->h_st.executed;
    if (snapshots) *snapshots = (int)c->run_snap_k;
    g_err.clear();
    return HB_OK;
}

// A completed run rebuilt from saved results (the reference's get_or_compute reopens
// its counters and accumulator files, hyper_ball.rs:165-201): the registers go to both
// buffers (the reference's counters file is the one it reads), the accumulators as
// they are, each permuted into the device's node order; the estimator tables are
// staged as hb_run stages them.  The restored run reports 0 iterations, as the
// reference's restored struct does (iters: 0).
extern "C" int hb_restore(hb_ctx *c, int p, const uint8_t *regs, const double *dist, const double *inv)
{
    if (!c || (!regs && c->n) || (!dist && c->n) || (!inv && c->n)) return fail(HB_EINVAL, "hb_restore: NULL argument");
    if (p < 4 || p > 16) return fail(HB_EINVAL, "hb_restore: precision p=%d outside 4..16", p);
    if (!c->off.ptr) return fail(HB_ESTATE, "hb_restore: no graph staged");
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    c->ran = false;
    const uint32_t n = (uint32_t)c->n, m = 1u << p;
    if (int rc = ensure_layout(c)) return rc;
    const size_t rbytes = (size_t)n << p;
    HB_TRY(c->regs0.ensure(std::max<size_t>(rbytes, 1)));
    HB_TRY(c->regs1.ensure(std::max<size_t>(rbytes, 1)));
    HB_TRY(c->dist.ensure(std::max<uint32_t>(n, 1)));
    HB_TRY(c->inv.ensure(std::max<uint32_t>(n, 1)));
    HB_TRY(c->cent.ensure(std::max<uint32_t>(n, 1)));
    HB_TRY(c->lc.ensure(m + 1));
    HB_TRY(c->bias.ensure(1));
    std::vector<float> lc(m + 1, 0.0f);
    for (uint32_t z = 1; z <= m; ++z) lc[z] = (float)((double)m * std::log((double)m / (double)z));
    HB_TRY(hipMemcpyAsync(c->lc.ptr, lc.data(), (m + 1) * 4, hipMemcpyHostToDevice, c->stream));
    HB_TRY(hipMemcpyAsync(c->bias.ptr, hb::bias_table(p), sizeof(HbBiasTable), hipMemcpyHostToDevice, c->stream));
    if (n) {
        if (c->relabelled()) {  // device row i = the caller's row perm[i]
            HB_TRY(hipMemcpyAsync(c->regs1.ptr, regs, rbytes, hipMemcpyHostToDevice, c->stream));
            HB_TRY(hb::unpermute_rows(c->regs1.ptr, c->regs0.ptr, c->perm.ptr, n, m, c->stream));
            HB_TRY(hipMemcpyAsync(c->cent.ptr, dist, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            HB_TRY(hb::unpermute_rows(c->cent.ptr, c->dist.ptr, c->perm.ptr, n, 8, c->stream));
            HB_TRY(hipMemcpyAsync(c->cent.ptr, inv, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            HB_TRY(hb::unpermute_rows(c->cent.ptr, c->inv.ptr, c->perm.ptr, n, 8, c->stream));
        } else {
            HB_TRY(hipMemcpyAsync(c->regs0.ptr, regs, rbytes, hipMemcpyHostToDevice, c->stream));
            HB_TRY(hipMemcpyAsync(c->dist.ptr, dist, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            HB_TRY(hipMemcpyAsync(c->inv.ptr, inv, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        }
        HB_TRY(hipMemcpyAsync(c->regs1.ptr, c->regs0.ptr, rbytes, hipMemcpyDeviceToDevice, c->stream));
    }
    HB_TRY(hipStreamSynchronize(c->stream));
    HbDev &d = c->dev;
    d = HbDev{};
    d.off = c->run_off();
    d.nbr = c->run_nbr();
    d.orig = c->run_orig();
    d.node_id = c->run_node_id();
    d.regs0 = c->regs0.ptr;
    d.regs1 = c->regs1.ptr;
    d.dist = c->dist.ptr;
    d.inv = c->inv.ptr;
    d.lc = c->lc.ptr;
    d.bias = c->bias.ptr;
    d.n = n;
    d.p = (uint32_t)p;
    d.am = alpha_f32(m) * ((float)m * (float)m);
    d.five_m = 5.0f * (float)m;
    d.lc_thr = kLcThreshold[p];
    c->h_st = HbIterState{};
    c->h_st.parity = 1;  // HB_REGS_LATEST selects regs0 as well
    c->h_st.converged = 1;
    c->h_rec.assign(1, HbIterRec{});
    c->p = p;
    c->run_snap_k = 0;
    c->ran = true;
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_set_snapshots(hb_ctx *c, int k)
{
    if (!c) return fail(HB_EINVAL, "hb_set_snapshots: ctx is NULL");
    if (k < 0 || k > 16) return fail(HB_EINVAL, "hb_set_snapshots: k=%d outside 0..16", k);
    c->snap_k = (uint32_t)k;
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_read_snapshots(hb_ctx *c, float *dst)
{
    if (!c || !dst) return fail(HB_EINVAL, "hb_read_snapshots: NULL argument");
    // rows of the completed run: hb_set_snapshots after it changes the next run only
    if (!c->ran || !c->run_snap_k || !c->snaps.ptr || c->snaps.count < (size_t)c->run_snap_k * c->n)
        return fail(HB_ESTATE, "hb_read_snapshots: no run with snapshots");
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    for (uint32_t t = 0; t < c->run_snap_k; ++t)
        if (int rc = copy_out(c, c->snaps.ptr + (size_t)t * c->n, dst + (size_t)t * c->n, 4)) return rc;
    g_err.clear();
    return HB_OK;
}

// nodes the surrogate network runs on: in the degree layout the nodes without a
// neighbour are the last ones, and their output is fixed without it
static uint32_t infer_nodes(const hb_ctx *c)
{
    if (!c->relabelled()) return (uint32_t)c->n;
    const std::vector<uint32_t> &ho = c->run_hoff();
    uint32_t lo = 0, hi = (uint32_t)c->n;  // first node of degree 0 (degrees descend)
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (ho[mid + 1] > ho[mid]) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

extern "C" int hb_surrogate(hb_ctx *c, hb_mlp *mlp, int p_feat, float *pred, float *features, double *ms)
{
    if (!c || !mlp) return fail(HB_EINVAL, "hb_surrogate: NULL argument");
    if (!c->off.ptr || c->n == 0) return fail(HB_ESTATE, "hb_surrogate: no graph staged");
    if (hb::mlp_device(mlp) != c->device) return fail(HB_EINVAL, "hb_surrogate: model and graph on different devices");
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    const uint32_t keep = c->snap_k;
    c->snap_k = 4;
    auto st = std::make_unique<hb_stats>();
    int rc = hb_run(c, p_feat, 4, 0, st.get());
    c->snap_k = keep;
    if (rc != HB_OK) return rc;
    // device time: register init + the 4 iterations (hb_run's events, host set-up excluded)
    // + features and inference
    HB_TRY(c->feat.ensure(c->n * 6));
    HB_TRY(c->pred.ensure(c->n));
    hipEvent_t e0, e1;
    HB_TRY(hipEventCreate(&e0));
    HB_TRY(hipEventCreate(&e1));
    HB_TRY(hipEventRecord(e0, c->stream));
    HB_TRY(hb::launch_surrogate_tail(mlp, c->run_off(), c->snaps.ptr, (uint32_t)c->n, infer_nodes(c),
                                     features != nullptr, c->feat.ptr, c->pred.ptr, c->stream));
    HB_TRY(hipEventRecord(e1, c->stream));
    HB_TRY(hipEventSynchronize(e1));
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = st->ms_init + st->ms_total + t;
    if (pred)
        if (int rc2 = copy_out(c, c->pred.ptr, pred, 4)) return rc2;
    if (features)
        if (int rc2 = copy_out(c, c->feat.ptr, features, 24)) return rc2;
    g_err.clear();
    return HB_OK;
}

// walk-count inputs of every node into c->feat (device numbering): the hub chunk table
// comes from the host offsets; *ms = device time of the passes and the inputs, plus
// standardisation and inference when mlp is given (then c->pred holds the ranking)
// want_x: the inputs of every node are read back (raw, or standardised when mlp is given)
static int walk_surrogate(hb_ctx *c, hb_mlp *mlp, int passes, bool want_x, float *ms)
{
    const uint32_t n = (uint32_t)c->n;
    if (!c->walk_hub.ptr) {  // built once per staged graph and numbering (release_run drops it)
        const std::vector<uint32_t> &ho = c->run_hoff();
        std::vector<uint32_t> beg, end, node, first;
        for (uint32_t u = 0; u < n; ++u) {
            if (ho[u + 1] - ho[u] <= hb::kWalkHub) continue;
            node.push_back(u);
            first.push_back((uint32_t)beg.size());
            for (uint64_t b = ho[u]; b < ho[u + 1]; b += hb::kWalkHub) {
                beg.push_back((uint32_t)b);
                end.push_back((uint32_t)std::min<uint64_t>(b + hb::kWalkHub, ho[u + 1]));
            }
        }
        first.push_back((uint32_t)beg.size());
        std::vector<uint32_t> table;
        table.reserve(2 * beg.size() + 2 * node.size() + 1);
        table.insert(table.end(), beg.begin(), beg.end());
        table.insert(table.end(), end.begin(), end.end());
        table.insert(table.end(), node.begin(), node.end());
        table.insert(table.end(), first.begin(), first.end());
        HB_TRY(c->walk_hub.ensure(table.size() + n));
        HB_TRY(hipMemcpyAsync(c->walk_hub.ptr, table.data(), table.size() * 4, hipMemcpyHostToDevice, c->stream));
        HB_TRY(hipStreamSynchronize(c->stream));  // `table` is a pageable host buffer
        c->walk_chunks = (uint32_t)beg.size();
        c->walk_hubs = (uint32_t)node.size();
    }
    hb::WalkHubs hubs;
    hubs.chunks = c->walk_chunks;
    hubs.hubs = c->walk_hubs;
    hubs.beg = c->walk_hub.ptr;
    hubs.end = hubs.beg + hubs.chunks;
    hubs.node = hubs.end + hubs.chunks;
    hubs.first = hubs.node + hubs.hubs;
    uint32_t *deg = c->walk_hub.ptr + 2 * (size_t)hubs.chunks + 2 * (size_t)hubs.hubs + 1;
    HB_TRY(c->walk.ensure(4 * (size_t)n + 2 * (size_t)hubs.chunks));
    HB_TRY(c->feat.ensure((size_t)n * 6));
    if (mlp) HB_TRY(c->pred.ensure(n));
    hipEvent_t e0, e1;
    HB_TRY(hipEventCreate(&e0));
    HB_TRY(hipEventCreate(&e1));
    hipError_t e = hipEventRecord(e0, c->stream);
    if (e == hipSuccess)
        // without a read-back only the nodes the network runs on need inputs (the
        // degree layout puts the nodes without a neighbour last)
        e = hb::walk_inputs(c->run_off(), c->run_nbr(), n, hubs, passes, c->walk.ptr, deg,
                            c->walk.ptr + 4 * (size_t)n, c->feat.ptr, want_x || !mlp ? n : infer_nodes(c),
                            c->stream);
    if (e == hipSuccess && mlp)
        e = hb::launch_surrogate_infer(mlp, c->run_off(), n, infer_nodes(c), want_x, c->feat.ptr, c->pred.ptr,
                                       c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.0f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&t, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(HB_EHIP, "walk-count surrogate: %s", hipGetErrorString(e));
    if (ms) *ms = t;
    return HB_OK;
}

extern "C" int hb_walk_features(hb_ctx *c, int passes, float *features, double *ms)
{
    if (!c) return fail(HB_EINVAL, "hb_walk_features: NULL argument");
    if (passes != 2 && passes != 3) return fail(HB_EINVAL, "hb_walk_features: passes=%d (2 or 3)", passes);
    if (!c->off.ptr) return fail(HB_ESTATE, "hb_walk_features: no graph staged");
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    if (c->n == 0) {
        if (ms) *ms = 0.0;
        return HB_OK;
    }
    float t = 0.0f;
    if (int rc = walk_surrogate(c, nullptr, passes, features != nullptr, &t)) return rc;
    if (ms) *ms = t;
    if (features)
        if (int rc = copy_out(c, c->feat.ptr, features, 24)) return rc;
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_surrogate_walks(hb_ctx *c, hb_mlp *mlp, int passes, float *pred, float *features, double *ms)
{
    if (!c || !mlp) return fail(HB_EINVAL, "hb_surrogate_walks: NULL argument");
    if (passes != 2 && passes != 3) return fail(HB_EINVAL, "hb_surrogate_walks: passes=%d (2 or 3)", passes);
    if (!c->off.ptr || c->n == 0) return fail(HB_ESTATE, "hb_surrogate_walks: no graph staged");
    if (hb::mlp_device(mlp) != c->device)
        return fail(HB_EINVAL, "hb_surrogate_walks: model and graph on different devices");
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    float t = 0.0f;
    if (int rc = walk_surrogate(c, mlp, passes, features != nullptr, &t)) return rc;
    if (ms) *ms = t;
    if (pred)
        if (int rc = copy_out(c, c->pred.ptr, pred, 4)) return rc;
    if (features)
        if (int rc = copy_out(c, c->feat.ptr, features, 24)) return rc;
    g_err.clear();
    return HB_OK;
}

// ---------------------------------------------------------------------------
// graph ingestion on the device (hb_ingest.hip)
// ---------------------------------------------------------------------------
static const char *mtx_reason(int code)
{
    switch (code) {
    case HB_MTX_BAD_HEADER: return "first line does not start with %%MatrixMarket";
    case HB_MTX_MALFORMED_HEADER: return "header has fewer than 5 tokens";
    case HB_MTX_UNSUPPORTED: return "only 'matrix coordinate pattern|integer symmetric|skew-symmetric|general'";
    case HB_MTX_NO_SIZE_LINE: return "no size line";
    case HB_MTX_BAD_SIZE_LINE: return "size line needs nrows ncols nnz";
    case HB_MTX_NOT_SQUARE: return "nrows != ncols";
    case HB_MTX_BAD_ENTRY: return "index is not a non-negative integer";
    case HB_MTX_ZERO_INDEX: return "index 0 (MatrixMarket indices are 1-based)";
    case HB_MTX_OUT_OF_BOUNDS: return "index above n";
    default: return "error";
    }
}

// adopts a device-built CSR and finishes staging
static int stage_built(hb_ctx *c, uint64_t n, hb::CsrBuild &b)
{
    c->off.adopt(b.off, n + 1);
    c->nbr.adopt(b.nbr, std::max<uint64_t>(b.arcs, 1));
    b.off = b.nbr = nullptr;
    return finish_stage(c, n, b.arcs, nullptr);
}

static void free_build(hb::CsrBuild &b)
{
    if (b.off) (void)hipFree(b.off);
    if (b.nbr) (void)hipFree(b.nbr);
    b.off = b.nbr = nullptr;
}

extern "C" int hb_stage_mtx(hb_ctx *c, const char *text, uint64_t len, hb_mtx_info *info)
{
    hb_mtx_info local{};
    if (!info) info = &local;
    memset(info, 0, sizeof *info);
    if (!c || (!text && len)) return fail(HB_EINVAL, "hb_stage_mtx: NULL argument");
    hb::ingest::MtxHeader h;
    const int code = hb::ingest::mtx_header(text ? text : "", len, &h);
    info->n = h.n;
    info->nnz = h.nnz;
    info->symmetric = h.symmetric;
    if (code) {
        info->err_code = code;
        info->err_line = code <= HB_MTX_NO_SIZE_LINE ? (code == HB_MTX_NO_SIZE_LINE ? 0 : 1) : h.err_line;
        return fail(HB_EINVAL, "hb_stage_mtx: line %llu: %s", (unsigned long long)info->err_line, mtx_reason(code));
    }
    if (h.n >= (1ull << 32) - 1)
        return fail(HB_ERANGE, "hb_stage_mtx: n=%llu exceeds the u32 device layout", (unsigned long long)h.n);
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    (void)hipStreamSynchronize(c->stream);
    c->release_run();
    c->release_layout();
    c->n = c->arcs = 0;
    const uint64_t body = len - h.body;
    uint8_t *dtext = nullptr;
    HB_TRY(hipMalloc(&dtext, std::max<uint64_t>(body, 1)));
    hipError_t e = body ? hipMemcpyAsync(dtext, text + h.body, body, hipMemcpyDefault, c->stream) : hipSuccess;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
    hb::CsrBuild b{};
    if (e == hipSuccess) e = hb::ingest::mtx_body_to_csr(dtext, body, (uint32_t)h.n, h.symmetric != 0, &b, c->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(dtext);
    info->ms_device = ms;
    if (e != hipSuccess) {
        free_build(b);
        if (b.emitted >= (1ull << 32))
            return fail(HB_ERANGE, "hb_stage_mtx: %llu arcs exceed the u32 device layout",
                        (unsigned long long)b.emitted);
        return fail(e == hipErrorOutOfMemory ? HB_ENOMEM : HB_EHIP, "hb_stage_mtx: %s", hipGetErrorString(e));
    }
    if (b.err_code) {
        free_build(b);
        info->err_code = b.err_code;
        info->err_line = h.body_line + b.err_index;
        return fail(HB_EINVAL, "hb_stage_mtx: line %llu: %s", (unsigned long long)info->err_line,
                    mtx_reason(b.err_code));
    }
    info->entries = b.entries;
    info->arcs = b.arcs;
    if (int rc = stage_built(c, h.n, b)) return rc;
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_stage_coo(hb_ctx *c, uint64_t n, const uint32_t *src, const uint32_t *dst, uint64_t m,
                            uint32_t flags)
{
    if (!c || (m && (!src || !dst))) return fail(HB_EINVAL, "hb_stage_coo: NULL argument");
    if (flags & ~(HB_COO_SYMMETRIZE | HB_COO_DEDUP | HB_COO_NO_LOOPS))
        return fail(HB_EINVAL, "hb_stage_coo: unknown flags 0x%x", flags);
    if (n >= (1ull << 32) - 1 || m >= (1ull << 32))
        return fail(HB_ERANGE, "hb_stage_coo: n=%llu m=%llu exceed the u32 device layout", (unsigned long long)n,
                    (unsigned long long)m);
    if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
    (void)hipStreamSynchronize(c->stream);
    c->release_run();
    c->release_layout();
    c->n = c->arcs = 0;
    uint32_t *ds = nullptr, *dd = nullptr;
    hipError_t e = hipMalloc(&ds, std::max<uint64_t>(m, 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&dd, std::max<uint64_t>(m, 1) * 4);
    if (e == hipSuccess && m) e = hipMemcpyAsync(ds, src, m * 4, hipMemcpyDefault, c->stream);
    if (e == hipSuccess && m) e = hipMemcpyAsync(dd, dst, m * 4, hipMemcpyDefault, c->stream);
    hb::CsrBuild b{};
    if (e == hipSuccess) e = hb::ingest::coo_to_csr(ds, dd, m, (uint32_t)n, flags, &b, c->stream);
    (void)hipStreamSynchronize(c->stream);
    if (ds) (void)hipFree(ds);
    if (dd) (void)hipFree(dd);
    if (e != hipSuccess) {
        free_build(b);
        if (b.emitted >= (1ull << 32))
            return fail(HB_ERANGE, "hb_stage_coo: %llu arcs exceed the u32 device layout",
                        (unsigned long long)b.emitted);
        return fail(e == hipErrorOutOfMemory ? HB_ENOMEM : HB_EHIP, "hb_stage_coo: %s", hipGetErrorString(e));
    }
    if (b.err_code) {
        free_build(b);
        return fail(HB_EINVAL, "hb_stage_coo: entry %llu has an id >= n=%llu", (unsigned long long)b.err_index,
                    (unsigned long long)n);
    }
    if (int rc = stage_built(c, n, b)) return rc;
    g_err.clear();
    return HB_OK;
}

extern "C" int hb_read_csr(hb_ctx *c, uint64_t *offsets, uint64_t *nbrs)
{
    if (!c) return fail(HB_EINVAL, "hb_read_csr: ctx is NULL");
    if (!c->off.ptr) return fail(HB_ESTATE, "hb_read_csr: no graph staged");
    if (offsets)
        for (uint64_t i = 0; i <= c->n; ++i) offsets[i] = c->h_off[i];
    if (nbrs && c->arcs) {
        if (!set_device(c)) return fail(HB_EHIP, "hipSetDevice failed");
        std::vector<uint32_t> tmp(c->arcs);
        HB_TRY(hipMemcpy(tmp.data(), c->nbr.ptr, c->arcs * 4, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < c->arcs; ++i) nbrs[i] = tmp[i];
    }
    g_err.clear();
    return HB_OK;
}

// ---------------------------------------------------------------------------
// many graphs at once: every context runs the single-graph engine on its own stream,
// driven by a pool of host threads, so the GPU interleaves the graphs' kernels
// ---------------------------------------------------------------------------
extern "C" int hb_run_many(hb_ctx **ctxs, const int *p, int count, int max_iters, uint32_t flags,
                           hb_stats *stats, int threads)
{
    if (count < 0 || (count && (!ctxs || !p))) return fail(HB_EINVAL, "hb_run_many: bad arguments");
    for (int i = 0; i < count; ++i)
        if (!ctxs[i]) return fail(HB_EINVAL, "hb_run_many: context %d is NULL", i);
    if (threads <= 0) threads = std::min(count, 8);
    threads = std::max(1, std::min(threads, count));
    std::vector<int> rc(count, HB_OK);
    std::vector<std::string> msg(count);
    std::atomic<int> next{0};
    auto work = [&]() {
        for (int i = next++; i < count; i = next++) {
            rc[i] = hb_run(ctxs[i], p[i], max_iters, flags, stats ? stats + i : nullptr);
            if (rc[i] != HB_OK) msg[i] = g_err;  // this worker's thread-local message
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    for (int i = 0; i < count; ++i)
        if (rc[i] != HB_OK) return fail(rc[i], "hb_run_many: context %d: %s", i, msg[i].c_str());
    g_err.clear();
    return HB_OK;
}

// The estimator's bias windows of precision p as the kernels read them (host-built;
// no device work), for checking against the k-nearest-neighbour rule on the CPU.
extern "C" int hb_bias_windows(int p, double *mean, float *breakpoints, uint8_t *buckets)
{
    if (p < HB_BIAS_PMIN || p > HB_BIAS_PMAX || !mean || !breakpoints || !buckets)
        return fail(HB_EINVAL, "hb_bias_windows: p=%d not in %d..%d or NULL argument", p, HB_BIAS_PMIN,
                    HB_BIAS_PMAX);
    const HbBiasTable *t = hb::bias_table(p);
    memcpy(mean, t->mean, HB_BIAS_WINDOWS * sizeof(double));
    memcpy(breakpoints, t->bp, HB_BIAS_WINDOWS * sizeof(float));
    memcpy(buckets, t->bucket, HB_BIAS_BUCKETS);
    g_err.clear();
    return HB_OK;
}
