// hb_eval.hip -- the reference's evaluation metrics on the GPU (src/utils.rs:57-100):
// MAE, MAPE (terms with b == 0 count 0), ranks from a stable descending sort (ties
// keep index order, no averaging) and Spearman rho = 1 - 6 ssd / (n max(n^2 - 1, 1)),
// plus the top-k overlap used to score anytime rankings.
//
// Ranks (index work) are bit-identical to the reference's: rocPRIM's stable
// descending radix sort of (value, index) pairs, with -0.0 mapped to +0.0 first
// (partial_cmp holds them equal).  ssd is summed exactly in 128-bit integers and
// rounded once; the reference's sequential f64 sum can differ from it by a few ulps
// once partial sums pass 2^53.  The MAE / MAPE sums are fixed-order two-stage f64
// reductions (deterministic for a given n), within a few ulps per term of the
// reference's sequential sums.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include "../../include/sparkx_hb.h"
#include "hb_kernels.h"

namespace hb {
namespace eval {

constexpr int kBlock = 256;
constexpr unsigned kSumBlocks = 1024;  // partials of the two-stage reductions

struct Partial {
    double abs_err, rel_err;
    unsigned long long ssd_lo, ssd_hi;
};

// sort keys with -0.0 mapped to +0.0, the identity permutation, and the lowest index
// of a NaN (if any) in *nan_at
__global__ __launch_bounds__(kBlock) void k_prep(const double *x, uint64_t n, double *key, uint32_t *idx,
                                                 unsigned long long *nan_at)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double v = x[i];
        if (v != v) atomicMin(nan_at, (unsigned long long)i);
        key[i] = v == 0.0 ? 0.0 : v;
        idx[i] = (uint32_t)i;
    }
}

__global__ __launch_bounds__(kBlock) void k_ranks(const uint32_t *order, uint64_t n, uint32_t *rank)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (uint64_t)gridDim.x * blockDim.x)
        rank[order[k]] = (uint32_t)k;
}

__device__ inline void add128(unsigned long long &lo, unsigned long long &hi, unsigned long long xlo,
                              unsigned long long xhi)
{
    const unsigned long long s = lo + xlo;
    hi += xhi + (s < lo);
    lo = s;
}

// per block: sum |a - b|, sum (b == 0 ? 0 : |a - b| / |b|) and the exact 128-bit
// sum of (rank_a - rank_b)^2, over a grid-stride slice, reduced in a fixed tree
__global__ __launch_bounds__(kBlock) void k_terms(const double *a, const double *b, const uint32_t *ra,
                                                  const uint32_t *rb, uint64_t n, Partial *part)
{
    __shared__ Partial red[kBlock];
    Partial p{0.0, 0.0, 0, 0};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const double y = b[i];
        const double e = fabs(a[i] - y);
        p.abs_err += e;
        p.rel_err += y == 0.0 ? 0.0 : e / fabs(y);
        const long long d = (long long)ra[i] - (long long)rb[i];
        const unsigned long long ad = (unsigned long long)(d < 0 ? -d : d);
        add128(p.ssd_lo, p.ssd_hi, ad * ad, 0);  // ad < 2^32, so ad^2 < 2^64
    }
    red[threadIdx.x] = p;
    __syncthreads();
    for (int w = kBlock / 2; w; w >>= 1) {
        if ((int)threadIdx.x < w) {
            Partial &q = red[threadIdx.x];
            const Partial &r = red[threadIdx.x + w];
            q.abs_err += r.abs_err;
            q.rel_err += r.rel_err;
            add128(q.ssd_lo, q.ssd_hi, r.ssd_lo, r.ssd_hi);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kBlock) void k_top_mark(const uint32_t *order, uint64_t k, uint8_t *flag)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x)
        flag[order[i]] = 1;
}

// per block: how many of order[0..k) are flagged
__global__ __launch_bounds__(kBlock) void k_top_count(const uint32_t *order, uint64_t k, const uint8_t *flag,
                                                      unsigned long long *part)
{
    __shared__ unsigned long long red[kBlock];
    unsigned long long c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += (uint64_t)gridDim.x * blockDim.x)
        c += flag[order[i]];
    red[threadIdx.x] = c;
    __syncthreads();
    for (int w = kBlock / 2; w; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace eval
}  // namespace hb

namespace {
int evfail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    hb::set_error(buf);
    return code;
}

struct DevMem {
    void *p = nullptr;
    ~DevMem()
    {
        if (p) (void)hipFree(p);
    }
};

struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard()
    {
        if (s) (void)hipStreamDestroy(s);
    }
};

#define HBE_TRY(call)                                                                                      \
    do {                                                                                                   \
        const hipError_t e_ = (call);                                                                      \
        if (e_ != hipSuccess)                                                                              \
            return evfail(e_ == hipErrorOutOfMemory ? HB_ENOMEM : HB_EHIP, "%s: %s: %s", fn, #call,        \
                          hipGetErrorString(e_));                                                          \
    } while (0)

int check_args(const char *fn, int device, uint64_t n, const double *const *vs, int nv)
{
    if (n == 0 || n >= (1ull << 32))
        return evfail(HB_EINVAL, "%s: n=%llu not in 1..2^32-1", fn, (unsigned long long)n);
    for (int j = 0; j < nv; ++j)
        if (!vs[j]) return evfail(HB_EINVAL, "%s: NULL argument", fn);
    int ndev = 0;
    HBE_TRY(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return evfail(HB_EINVAL, "%s: bad device %d", fn, device);
    HBE_TRY(hipSetDevice(device));
    return HB_OK;
}

// Sorts device values (stable, descending) into order[] and, optionally, rank[];
// nan_at[which] receives the lowest index of a NaN.
struct Ranker {
    const char *fn;
    uint64_t n;
    hipStream_t s;
    DevMem kin, kout, iin, tmp, nan;
    size_t tmp_bytes = 0;
    unsigned grid = 1;

    int init(int arrays)
    {
        HBE_TRY(hipMalloc(&kin.p, n * 8));
        HBE_TRY(hipMalloc(&kout.p, n * 8));
        HBE_TRY(hipMalloc(&iin.p, n * 4));
        HBE_TRY(hipMalloc(&nan.p, arrays * 8));
        HBE_TRY(hipMemsetAsync(nan.p, 0xff, arrays * 8, s));
        HBE_TRY(rocprim::radix_sort_pairs_desc(nullptr, tmp_bytes, (double *)kin.p, (double *)kout.p,
                                               (uint32_t *)iin.p, (uint32_t *)iin.p, (size_t)n, 0, 64, s));
        HBE_TRY(hipMalloc(&tmp.p, std::max<size_t>(tmp_bytes, 1)));
        grid = (unsigned)std::min<uint64_t>((n + hb::eval::kBlock - 1) / hb::eval::kBlock, 8192);
        return HB_OK;
    }
    int rank(int which, const double *d_values, uint32_t *order, uint32_t *rank_out)
    {
        using namespace hb::eval;
        k_prep<<<grid, kBlock, 0, s>>>(d_values, n, (double *)kin.p, (uint32_t *)iin.p,
                                       (unsigned long long *)nan.p + which);
        HBE_TRY(hipGetLastError());
        HBE_TRY(rocprim::radix_sort_pairs_desc(tmp.p, tmp_bytes, (double *)kin.p, (double *)kout.p,
                                               (uint32_t *)iin.p, order, (size_t)n, 0, 64, s));
        if (rank_out) {
            k_ranks<<<grid, kBlock, 0, s>>>(order, n, rank_out);
            HBE_TRY(hipGetLastError());
        }
        return HB_OK;
    }
    // after the stream has synchronised
    int check_nan(int arrays)
    {
        unsigned long long at[2] = {~0ull, ~0ull};
        HBE_TRY(hipMemcpy(at, nan.p, arrays * 8, hipMemcpyDeviceToHost));
        for (int j = 0; j < arrays; ++j)
            if (at[j] != ~0ull)
                return evfail(HB_EINVAL, "%s: NaN at index %llu of argument %d (the reference's order is "
                              "unspecified for it)", fn, at[j], j + 1);
        return HB_OK;
    }
};

double u128_to_double(unsigned long long lo, unsigned long long hi)
{
    return (double)(((unsigned __int128)hi << 64) | lo);  // one round-to-nearest-even
}
}  // namespace

extern "C" int hb_eval_ranks(int device, const double *v, uint64_t n, uint32_t *rank)
{
    static const char fn[] = "hb_eval_ranks";
    const double *vs[1] = {v};
    if (!rank) return evfail(HB_EINVAL, "%s: NULL argument", fn);
    if (int rc = check_args(fn, device, n, vs, 1)) return rc;
    StreamGuard sg;
    HBE_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
    DevMem dv, order, rk;
    HBE_TRY(hipMalloc(&dv.p, n * 8));
    HBE_TRY(hipMalloc(&order.p, n * 4));
    HBE_TRY(hipMalloc(&rk.p, n * 4));
    HBE_TRY(hipMemcpyAsync(dv.p, v, n * 8, hipMemcpyDefault, sg.s));
    Ranker r{fn, n, sg.s};
    if (int rc = r.init(1)) return rc;
    if (int rc = r.rank(0, (const double *)dv.p, (uint32_t *)order.p, (uint32_t *)rk.p)) return rc;
    HBE_TRY(hipMemcpyAsync(rank, rk.p, n * 4, hipMemcpyDefault, sg.s));
    HBE_TRY(hipStreamSynchronize(sg.s));
    if (int rc = r.check_nan(1)) return rc;
    hb::set_error("");
    return HB_OK;
}

extern "C" int hb_eval_metrics(int device, const double *a, const double *b, uint64_t n, double top_frac,
                               double *out)
{
    using namespace hb::eval;
    static const char fn[] = "hb_eval_metrics";
    const double *vs[2] = {a, b};
    if (!out) return evfail(HB_EINVAL, "%s: NULL argument", fn);
    if (!(top_frac > 0.0 && top_frac <= 1.0)) return evfail(HB_EINVAL, "%s: top_frac not in (0, 1]", fn);
    if (int rc = check_args(fn, device, n, vs, 2)) return rc;
    StreamGuard sg;
    HBE_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
    const hipStream_t s = sg.s;
    DevMem va, vb, oa, ob, rka, rkb, flag, part, tpart;
    HBE_TRY(hipMalloc(&va.p, n * 8));
    HBE_TRY(hipMalloc(&vb.p, n * 8));
    HBE_TRY(hipMalloc(&oa.p, n * 4));
    HBE_TRY(hipMalloc(&ob.p, n * 4));
    HBE_TRY(hipMalloc(&rka.p, n * 4));
    HBE_TRY(hipMalloc(&rkb.p, n * 4));
    HBE_TRY(hipMalloc(&flag.p, n));
    HBE_TRY(hipMalloc(&part.p, kSumBlocks * sizeof(Partial)));
    HBE_TRY(hipMalloc(&tpart.p, kSumBlocks * 8));
    HBE_TRY(hipMemcpyAsync(va.p, a, n * 8, hipMemcpyDefault, s));
    HBE_TRY(hipMemcpyAsync(vb.p, b, n * 8, hipMemcpyDefault, s));
    Ranker r{fn, n, s};
    if (int rc = r.init(2)) return rc;
    if (int rc = r.rank(0, (const double *)va.p, (uint32_t *)oa.p, (uint32_t *)rka.p)) return rc;
    if (int rc = r.rank(1, (const double *)vb.p, (uint32_t *)ob.p, (uint32_t *)rkb.p)) return rc;
    k_terms<<<kSumBlocks, kBlock, 0, s>>>((const double *)va.p, (const double *)vb.p, (const uint32_t *)rka.p,
                                          (const uint32_t *)rkb.p, n, (Partial *)part.p);
    HBE_TRY(hipGetLastError());
    // top-k overlap: flag the top k of a, count flagged among the top k of b
    const uint64_t k = std::max<uint64_t>(1, (uint64_t)std::ceil(top_frac * (double)n));
    HBE_TRY(hipMemsetAsync(flag.p, 0, n, s));
    k_top_mark<<<r.grid, kBlock, 0, s>>>((const uint32_t *)oa.p, k, (uint8_t *)flag.p);
    HBE_TRY(hipGetLastError());
    k_top_count<<<kSumBlocks, kBlock, 0, s>>>((const uint32_t *)ob.p, k, (const uint8_t *)flag.p,
                                              (unsigned long long *)tpart.p);
    HBE_TRY(hipGetLastError());
    std::vector<Partial> hp(kSumBlocks);
    std::vector<unsigned long long> ht(kSumBlocks);
    HBE_TRY(hipMemcpyAsync(hp.data(), part.p, kSumBlocks * sizeof(Partial), hipMemcpyDeviceToHost, s));
    HBE_TRY(hipMemcpyAsync(ht.data(), tpart.p, kSumBlocks * 8, hipMemcpyDeviceToHost, s));
    HBE_TRY(hipStreamSynchronize(s));
    if (int rc = r.check_nan(2)) return rc;
    // second stage, in block order
    double s_abs = 0.0, s_rel = 0.0;
    unsigned long long lo = 0, hi = 0, common = 0;
    for (unsigned j = 0; j < kSumBlocks; ++j) {
        s_abs += hp[j].abs_err;
        s_rel += hp[j].rel_err;
        const unsigned long long t = lo + hp[j].ssd_lo;
        hi += hp[j].ssd_hi + (t < lo);
        lo = t;
        common += ht[j];
    }
    const double nf = (double)n;
    const double ssd = u128_to_double(lo, hi);
    out[0] = s_abs / nf;                                                 // utils.rs:57-59
    out[1] = 100.0 * s_rel / nf;                                         // utils.rs:60-74
    out[2] = 1.0 - (6.0 * ssd) / (nf * std::max(nf * nf - 1.0, 1.0));    // utils.rs:87-100
    out[3] = (double)common / (double)k;
    hb::set_error("");
    return HB_OK;
}

// ---------------------------------------------------------------------------
// hb_eval_metrics_noalloc: hb_eval_metrics for device arrays, with every
// intermediate buffer (sort keys and orders, ranks, top-k flags, block partials,
// the sort's temporary storage) carved out of one fixed device array instead of
// allocated per call.  The array is loaded with the library; each device holds
// its own copy at its own address, which host code obtains once per device with
// hipGetSymbolAddress and uses for every kernel argument and fill.  The array
// keeps the last call's contents, so everything that must start at zero or at
// "no NaN" is filled by a kernel at the start of each call; calls that overlap
// on one device share it and race.
// ---------------------------------------------------------------------------
static constexpr size_t kNoAllocScratchBytes = 32ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (set-up once per device)
int scratch_ptr(const char *fn, int device, unsigned char **out)
{
    if (device >= kMaxDevices) return evfail(HB_EINVAL, "%s: device %d above %d", fn, device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return evfail(HB_EHIP, "%s: hipGetSymbolAddress(scratch) on device %d: %s", fn, device,
                          hipGetErrorString(e));
        g_scratch_ptr[device] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[device];
    return HB_OK;
}

// a must be memory the device can read (device, managed or mapped host memory)
int check_device_ptr(const char *fn, const void *p, int arg)
{
    hipPointerAttribute_t at;
    memset(&at, 0, sizeof at);
    const hipError_t e = hipPointerGetAttributes(&at, p);
    (void)hipGetLastError();  // an unknown host pointer leaves an error behind
    if (e != hipSuccess || !at.devicePointer)
        return evfail(HB_EINVAL, "%s: argument %d is not device-accessible memory (this entry point takes "
                      "device arrays; hb_eval_metrics also takes host arrays)", fn, arg);
    return HB_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}  // namespace

extern "C" int hb_eval_metrics_noalloc(int device, const double *a, const double *b, uint64_t n, double top_frac,
                                       double *out)
{
    using namespace hb::eval;
    static const char fn[] = "hb_eval_metrics_noalloc";
    const double *vs[2] = {a, b};
    if (!out) return evfail(HB_EINVAL, "%s: NULL argument", fn);
    if (!(top_frac > 0.0 && top_frac <= 1.0)) return evfail(HB_EINVAL, "%s: top_frac not in (0, 1]", fn);
    if (int rc = check_args(fn, device, n, vs, 2)) return rc;
    if (int rc = check_device_ptr(fn, a, 1)) return rc;
    if (int rc = check_device_ptr(fn, b, 2)) return rc;
    unsigned char *base = nullptr;
    if (int rc = scratch_ptr(fn, device, &base)) return rc;
    const hipStream_t s = nullptr;  // the device's null stream; synchronised before returning
    // carve the fixed array for this n
    size_t at = 0;
    auto take = [&](size_t bytes) {
        unsigned char *p = base + at;
        at += align256(bytes);
        return p;
    };
    double *kin = (double *)take(n * 8), *kout = (double *)take(n * 8);
    uint32_t *iin = (uint32_t *)take(n * 4);
    uint32_t *oa = (uint32_t *)take(n * 4), *ob = (uint32_t *)take(n * 4);
    uint32_t *rka = (uint32_t *)take(n * 4), *rkb = (uint32_t *)take(n * 4);
    const uint64_t flag_words = (n + 3) / 4;
    uint8_t *flag = take(flag_words * 4);
    unsigned long long *nan = (unsigned long long *)take(2 * 8);
    Partial *part = (Partial *)take(kSumBlocks * sizeof(Partial));
    unsigned long long *tpart = (unsigned long long *)take(kSumBlocks * 8);
    size_t tmp_bytes = 0;
    HBE_TRY(rocprim::radix_sort_pairs_desc(nullptr, tmp_bytes, kin, kout, iin, oa, (size_t)n, 0, 64, s));
    void *tmp = take(std::max<size_t>(tmp_bytes, 1));
    if (at > kNoAllocScratchBytes)
        return evfail(HB_ERANGE, "%s: n=%llu needs %zu bytes of scratch; the fixed device array holds %zu "
                      "(n up to 2^19 fits; use hb_eval_metrics for larger inputs)", fn,
                      (unsigned long long)n, at, kNoAllocScratchBytes);
    // state left by the previous call: NaN slots to "none", top-k flags to zero
    HBE_TRY(hb::launch_fill((uint32_t *)nan, 4, 0xffffffffu, s));
    HBE_TRY(hb::launch_fill((uint32_t *)flag, flag_words, 0u, s));
    const unsigned grid = (unsigned)std::min<uint64_t>((n + kBlock - 1) / kBlock, 8192);
    const double *src[2] = {a, b};
    uint32_t *order[2] = {oa, ob}, *rank[2] = {rka, rkb};
    for (int j = 0; j < 2; ++j) {
        k_prep<<<grid, kBlock, 0, s>>>(src[j], n, kin, iin, nan + j);
        HBE_TRY(hipGetLastError());
        HBE_TRY(rocprim::radix_sort_pairs_desc(tmp, tmp_bytes, kin, kout, iin, order[j], (size_t)n, 0, 64, s));
        k_ranks<<<grid, kBlock, 0, s>>>(order[j], n, rank[j]);
        HBE_TRY(hipGetLastError());
    }
    k_terms<<<kSumBlocks, kBlock, 0, s>>>(a, b, rka, rkb, n, part);
    HBE_TRY(hipGetLastError());
    const uint64_t k = std::max<uint64_t>(1, (uint64_t)std::ceil(top_frac * (double)n));
    k_top_mark<<<grid, kBlock, 0, s>>>(oa, k, flag);
    HBE_TRY(hipGetLastError());
    k_top_count<<<kSumBlocks, kBlock, 0, s>>>(ob, k, flag, tpart);
    HBE_TRY(hipGetLastError());
    // block partials and NaN slots back to the host (fixed-size host arrays)
    static thread_local Partial hp[kSumBlocks];
    static thread_local unsigned long long ht[kSumBlocks];
    unsigned long long nan_at[2];
    HBE_TRY(hipMemcpyAsync(hp, part, sizeof hp, hipMemcpyDeviceToHost, s));
    HBE_TRY(hipMemcpyAsync(ht, tpart, sizeof ht, hipMemcpyDeviceToHost, s));
    HBE_TRY(hipMemcpyAsync(nan_at, nan, sizeof nan_at, hipMemcpyDeviceToHost, s));
    HBE_TRY(hipStreamSynchronize(s));
    for (int j = 0; j < 2; ++j)
        if (nan_at[j] != ~0ull)
            return evfail(HB_EINVAL, "%s: NaN at index %llu of argument %d (the reference's order is "
                          "unspecified for it)", fn, nan_at[j], j + 1);
    double s_abs = 0.0, s_rel = 0.0;
    unsigned long long lo = 0, hi = 0, common = 0;
    for (unsigned j = 0; j < kSumBlocks; ++j) {  // second stage, in block order (as hb_eval_metrics)
        s_abs += hp[j].abs_err;
        s_rel += hp[j].rel_err;
        const unsigned long long t = lo + hp[j].ssd_lo;
        hi += hp[j].ssd_hi + (t < lo);
        lo = t;
        common += ht[j];
    }
    const double nf = (double)n;
    const double ssd = u128_to_double(lo, hi);
    out[0] = s_abs / nf;
    out[1] = 100.0 * s_rel / nf;
    out[2] = 1.0 - (6.0 * ssd) / (nf * std::max(nf * nf - 1.0, 1.0));
    out[3] = (double)common / (double)k;
    hb::set_error("");
    return HB_OK;
}
