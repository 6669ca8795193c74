// hb_ingest.hip -- graphs built on the device: MatrixMarket text, and COO edge lists
// with optional symmetrisation / deduplication / loop removal, into the u32 CSR the
// engine stages (offsets[n+1], each neighbour list ascending).
//
// MatrixMarket follows the reference's reader (src/graph/cache/parser_mtx/producers.rs,
// mod.rs; impl_sort.rs): the header and size line are read on the host; the body is
// split into lines on the device (positions after each '\n', by chunked count, scan
// and write passes), one thread parses each line (trim; blank and '%' lines skipped;
// "i j" 1-based, further tokens ignored, a line with one token skipped; optional '+';
// errors for a non-number, a 0 or an index above n), and every entry emits i->j plus
// j->i when the file is not "general" and i != j.  The reference then counts degrees,
// scatters and sorts each list; here the arcs are 64-bit keys (u << 32 | v) sorted once
// by rocPRIM's radix sort, which yields the same CSR: lists ascending, duplicates kept.
#include <algorithm>
#include <cstdint>
#include <cstring>

#include <hip/hip_runtime.h>
#include <rocprim/rocprim.hpp>

#include "hb_kernels.h"

namespace hb {
namespace ingest {
namespace {

constexpr int kBlock = 256;

unsigned grid_for(uint64_t items)
{
    return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((items + kBlock - 1) / kBlock, 65535));
}

__device__ __forceinline__ bool is_ws(uint8_t c)
{
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// Line splitting in two passes over chunks of kSplitChunk bytes, one block per chunk:
// the block reads the chunk 4 KiB at a time, 16 bytes per thread (coalesced), counts
// the '\n' of the chunk; after a scan of the counts the second pass re-reads the chunk
// the same way and writes the position after each '\n' (below len) in text order
// (a block-wide exclusive scan of the per-thread counts at every 4 KiB step).
constexpr uint32_t kSplitSteps = 16, kSplitChunk = kSplitSteps * kBlock * 16;

// bit b set when byte b of the 16 at t[p..p+16) is '\n' (bytes at or past len excluded)
__device__ __forceinline__ uint32_t nl_mask16(const uint8_t *t, uint64_t p, uint64_t len)
{
    uint32_t m = 0;
    if (p + 16 <= len) {
        const uint4 v = *(const uint4 *)(t + p);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int b = 0; b < 4; ++b) m |= (uint32_t)(((w[q] >> (8 * b)) & 0xffu) == '\n') << (4 * q + b);
    } else {
        for (uint64_t i = p; i < len; ++i) m |= (uint32_t)(t[i] == '\n') << (i - p);
    }
    return m;
}

// exclusive scan of x over the block; *total gets the block sum
__device__ __forceinline__ uint32_t block_exclusive(uint32_t x, uint32_t *lds, uint32_t *total)
{
    lds[threadIdx.x] = x;
    __syncthreads();
    for (int w = 1; w < kBlock; w <<= 1) {
        const uint32_t v = (int)threadIdx.x >= w ? lds[threadIdx.x - w] : 0u;
        __syncthreads();
        lds[threadIdx.x] += v;
        __syncthreads();
    }
    const uint32_t incl = lds[threadIdx.x];
    *total = lds[kBlock - 1];
    __syncthreads();
    return incl - x;
}

__global__ __launch_bounds__(kBlock) void k_count_nl(const uint8_t *t, uint64_t len, unsigned long long *cnt)
{
    __shared__ uint32_t red[kBlock];
    const uint64_t chunk = (uint64_t)blockIdx.x * kSplitChunk;
    uint32_t c = 0;
    for (uint32_t k = 0; k < kSplitSteps; ++k) {
        const uint64_t p = chunk + ((uint64_t)k * kBlock + threadIdx.x) * 16;
        if (p < len) c += __popc(nl_mask16(t, p, len));
    }
    red[threadIdx.x] = c;
    __syncthreads();
    for (int w = kBlock / 2; w; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) cnt[blockIdx.x] = red[0];
}

// starts[1 + base[chunk] + k] = position after the k-th '\n' of the chunk, if below len
__global__ __launch_bounds__(kBlock) void k_line_starts(const uint8_t *t, uint64_t len,
                                                        const unsigned long long *base, uint64_t *starts)
{
    __shared__ uint32_t lds[kBlock];
    const uint64_t chunk = (uint64_t)blockIdx.x * kSplitChunk;
    uint64_t run = 1 + base[blockIdx.x];
    for (uint32_t k = 0; k < kSplitSteps; ++k) {  // uniform trip count: barriers inside
        const uint64_t p = chunk + ((uint64_t)k * kBlock + threadIdx.x) * 16;
        uint32_t m = p < len ? nl_mask16(t, p, len) : 0u;
        uint32_t total = 0;
        uint64_t out = run + block_exclusive(__popc(m), lds, &total);
        while (m) {
            const uint32_t b = (uint32_t)(__ffs(m) - 1);
            m &= m - 1u;
            if (p + b + 1 < len) starts[out] = p + b + 1;
            ++out;
        }
        run += total;
    }
}

// usize token [b, e): optional '+', decimal digits, no overflow; false otherwise
__device__ bool parse_usize(const uint8_t *t, uint64_t b, uint64_t e, uint64_t *out)
{
    if (b < e && t[b] == '+') ++b;
    if (b >= e) return false;
    uint64_t v = 0;
    for (uint64_t i = b; i < e; ++i) {
        const uint32_t d = (uint32_t)t[i] - '0';
        if (d > 9) return false;
        if (v > (~0ull - d) / 10) return false;
        v = v * 10 + d;
    }
    *out = v;
    return true;
}

// per line: emits[l] in {0, 1, 2}; (i, j) 0-based; the first failing line in *err as
// (line << 4) | code, code 7 = not a number, 8 = zero index, 9 = index above n
__global__ __launch_bounds__(kBlock) void k_parse_lines(const uint8_t *t, const uint64_t *starts, uint64_t lines,
                                                        uint64_t n, int symmetric, uint32_t *emits, uint32_t *ei,
                                                        uint32_t *ej, unsigned long long *err)
{
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l < lines;
         l += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t b = starts[l], e = starts[l + 1];
        uint32_t out = 0;
        while (b < e && is_ws(t[b])) ++b;
        while (e > b && is_ws(t[e - 1])) --e;
        if (b < e && t[b] != '%') {
            uint64_t p = b;
            while (p < e && !is_ws(t[p])) ++p;
            uint64_t i1 = 0, j1 = 0;
            uint32_t code = 0;
            if (!parse_usize(t, b, p, &i1)) {
                code = 7;
            } else {
                while (p < e && is_ws(t[p])) ++p;
                const uint64_t q = p;
                while (p < e && !is_ws(t[p])) ++p;
                if (q < e) {  // a second token (else the line is skipped)
                    if (!parse_usize(t, q, p, &j1)) code = 7;
                    else if (i1 == 0 || j1 == 0) code = 8;
                    else if (i1 > n || j1 > n) code = 9;
                    else {
                        ei[l] = (uint32_t)(i1 - 1);
                        ej[l] = (uint32_t)(j1 - 1);
                        out = (symmetric && i1 != j1) ? 2u : 1u;
                    }
                }
            }
            if (code) atomicMin(err, ((unsigned long long)l << 4) | code);
        }
        emits[l] = out;
    }
}

// per COO entry: emits in {0, 1, 2} from the flags; out-of-range ids in *err as
// (index << 4) | 9
__global__ __launch_bounds__(kBlock) void k_coo_emits(const uint32_t *src, const uint32_t *dst, uint64_t m, uint64_t n,
                                                      uint32_t flags, uint32_t *emits, unsigned long long *err)
{
    for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t u = src[k], v = dst[k];
        uint32_t out = 0;
        if (u >= n || v >= n) atomicMin(err, ((unsigned long long)k << 4) | 9u);
        else if (u != v) out = (flags & HB_COO_SYMMETRIZE) ? 2u : 1u;
        else out = (flags & HB_COO_NO_LOOPS) ? 0u : 1u;
        emits[k] = out;
    }
}

__global__ __launch_bounds__(kBlock) void k_emit(const uint32_t *ei, const uint32_t *ej, const uint32_t *emits,
                                                 const unsigned long long *pos, uint64_t count,
                                                 unsigned long long *keys)
{
    for (uint64_t l = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; l < count;
         l += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t e = emits[l];
        if (!e) continue;
        const unsigned long long i = ei[l], j = ej[l], p = pos[l];
        keys[p] = (i << 32) | j;
        if (e == 2) keys[p + 1] = (j << 32) | i;
    }
}

// sorted keys -> offsets and neighbours; each arc fills the offsets of the nodes
// between its predecessor's tail and its own
__global__ __launch_bounds__(kBlock) void k_csr(const unsigned long long *keys, uint64_t m, uint32_t n, uint32_t *off,
                                                uint32_t *nbr)
{
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < m; p += (uint64_t)gridDim.x * blockDim.x) {
        const unsigned long long k = keys[p];
        const uint32_t u = (uint32_t)(k >> 32);
        nbr[p] = (uint32_t)k;
        const uint64_t first = p ? (uint64_t)(keys[p - 1] >> 32) + 1 : 0;
        for (uint64_t w = first; w <= u; ++w) off[w] = (uint32_t)p;
        if (p == m - 1)
            for (uint64_t w = (uint64_t)u + 1; w <= n; ++w) off[w] = (uint32_t)m;
    }
}

struct Tmp {
    void *p = nullptr;
    ~Tmp()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, std::max<size_t>(bytes, 1)); }
};

#define HBI_TRY(call)                    \
    do {                                 \
        const hipError_t e_ = (call);    \
        if (e_ != hipSuccess) return e_; \
    } while (0)

// emits (count entries) + endpoints -> sorted keys -> (optionally unique) -> CSR
hipError_t build_csr(const uint32_t *ei, const uint32_t *ej, const uint32_t *emits, uint64_t count, uint32_t n,
                     bool dedup, CsrBuild *out, hipStream_t s)
{
    Tmp pos, total, keys, keys2, tmp;
    if (count == 0) {  // nothing emitted: n empty lists
        out->emitted = out->arcs = 0;
        HBI_TRY(hipMalloc(&out->off, ((size_t)n + 1) * 4));
        HBI_TRY(hipMalloc(&out->nbr, 4));
        HBI_TRY(hipMemsetAsync(out->off, 0, ((size_t)n + 1) * 4, s));
        return hipStreamSynchronize(s);
    }
    HBI_TRY(pos.alloc(count * 8));
    HBI_TRY(total.alloc(8));
    size_t bytes = 0;
    auto wide = rocprim::make_transform_iterator(emits, [] __device__(uint32_t e) { return (unsigned long long)e; });
    HBI_TRY(rocprim::exclusive_scan(nullptr, bytes, wide, (unsigned long long *)pos.p, 0ull, (size_t)count,
                                    rocprim::plus<unsigned long long>(), s));
    HBI_TRY(tmp.alloc(bytes));
    HBI_TRY(rocprim::exclusive_scan(tmp.p, bytes, wide, (unsigned long long *)pos.p, 0ull, (size_t)count,
                                    rocprim::plus<unsigned long long>(), s));
    unsigned long long h_last[2] = {0, 0};
    if (count) {
        HBI_TRY(hipMemcpyAsync(&h_last[0], (unsigned long long *)pos.p + count - 1, 8, hipMemcpyDeviceToHost, s));
        uint32_t e_last = 0;
        HBI_TRY(hipMemcpyAsync(&e_last, emits + count - 1, 4, hipMemcpyDeviceToHost, s));
        HBI_TRY(hipStreamSynchronize(s));
        h_last[1] = h_last[0] + e_last;
    }
    uint64_t m = h_last[1];
    out->emitted = m;
    if (m >= (1ull << 32)) return hipErrorInvalidValue;  // reported as HB_ERANGE by the caller
    if (m == 0) {
        out->arcs = 0;
        HBI_TRY(hipMalloc(&out->off, ((size_t)n + 1) * 4));
        HBI_TRY(hipMalloc(&out->nbr, 4));
        HBI_TRY(hipMemsetAsync(out->off, 0, ((size_t)n + 1) * 4, s));
        return hipStreamSynchronize(s);
    }
    HBI_TRY(keys.alloc(m * 8));
    HBI_TRY(keys2.alloc(m * 8));
    k_emit<<<grid_for(count), kBlock, 0, s>>>(ei, ej, emits, (const unsigned long long *)pos.p, count,
                                             (unsigned long long *)keys.p);
    HBI_TRY(hipGetLastError());
    const unsigned bits = 32 + std::max(1u, 32u - (unsigned)__builtin_clz(std::max(n, 2u) - 1));
    size_t sort_bytes = 0;
    HBI_TRY(rocprim::radix_sort_keys(nullptr, sort_bytes, (unsigned long long *)keys.p,
                                     (unsigned long long *)keys2.p, (size_t)m, 0, bits, s));
    Tmp sort_tmp;
    HBI_TRY(sort_tmp.alloc(sort_bytes));
    HBI_TRY(rocprim::radix_sort_keys(sort_tmp.p, sort_bytes, (unsigned long long *)keys.p,
                                     (unsigned long long *)keys2.p, (size_t)m, 0, bits, s));
    const unsigned long long *sorted = (const unsigned long long *)keys2.p;
    if (dedup && m) {
        size_t ubytes = 0;
        HBI_TRY(rocprim::unique(nullptr, ubytes, sorted, (unsigned long long *)keys.p,
                                (unsigned long long *)total.p, (size_t)m, rocprim::equal_to<unsigned long long>(),
                                s));
        Tmp utmp;
        HBI_TRY(utmp.alloc(ubytes));
        HBI_TRY(rocprim::unique(utmp.p, ubytes, sorted, (unsigned long long *)keys.p,
                                (unsigned long long *)total.p, (size_t)m, rocprim::equal_to<unsigned long long>(),
                                s));
        unsigned long long mu = 0;
        HBI_TRY(hipMemcpyAsync(&mu, total.p, 8, hipMemcpyDeviceToHost, s));
        HBI_TRY(hipStreamSynchronize(s));
        m = mu;
        sorted = (const unsigned long long *)keys.p;
    }
    out->arcs = m;
    HBI_TRY(hipMalloc(&out->off, ((size_t)n + 1) * 4));
    HBI_TRY(hipMalloc(&out->nbr, std::max<size_t>(m, 1) * 4));
    if (m == 0) {
        HBI_TRY(hipMemsetAsync(out->off, 0, ((size_t)n + 1) * 4, s));
    } else {
        k_csr<<<grid_for(m), kBlock, 0, s>>>(sorted, m, n, out->off, out->nbr);
        HBI_TRY(hipGetLastError());
    }
    return hipStreamSynchronize(s);  // temporaries are freed on return
}

}  // namespace

// ---------------------------------------------------------------------------
// header and size line (producers.rs:15-75), on the host
// ---------------------------------------------------------------------------
namespace {
bool host_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; }

bool host_usize(const char *t, uint64_t b, uint64_t e, uint64_t *out)
{
    if (b < e && t[b] == '+') ++b;
    if (b >= e) return false;
    uint64_t v = 0;
    for (uint64_t i = b; i < e; ++i) {
        const unsigned d = (unsigned)(unsigned char)t[i] - '0';
        if (d > 9 || v > (~0ull - d) / 10) return false;
        v = v * 10 + d;
    }
    *out = v;
    return true;
}

bool word_ieq(const char *t, uint64_t b, uint64_t e, const char *w)
{
    const size_t n = strlen(w);
    if (e - b != n) return false;
    for (size_t i = 0; i < n; ++i) {
        char c = t[b + i];
        if (c >= 'A' && c <= 'Z') c = (char)(c - 'A' + 'a');
        if (c != w[i]) return false;
    }
    return true;
}

// tokens of [b, e) into (tb, te), at most cap; returns the count
int tokens(const char *t, uint64_t b, uint64_t e, uint64_t *tb, uint64_t *te, int cap)
{
    int k = 0;
    while (k < cap) {
        while (b < e && host_ws(t[b])) ++b;
        if (b >= e) break;
        tb[k] = b;
        while (b < e && !host_ws(t[b])) ++b;
        te[k++] = b;
    }
    return k;
}
}  // namespace

int mtx_header(const char *t, uint64_t len, MtxHeader *h)
{
    memset(h, 0, sizeof *h);
    const char *nl = (const char *)memchr(t, '\n', len);
    const uint64_t e0 = nl ? (uint64_t)(nl - t) : len;
    if (len < 14 || memcmp(t, "%%MatrixMarket", 14) != 0) return HB_MTX_BAD_HEADER;
    uint64_t tb[5], te[5];
    if (tokens(t, 0, e0, tb, te, 5) < 5) return HB_MTX_MALFORMED_HEADER;
    if (!word_ieq(t, tb[1], te[1], "matrix") || !word_ieq(t, tb[2], te[2], "coordinate") ||
        !(word_ieq(t, tb[3], te[3], "pattern") || word_ieq(t, tb[3], te[3], "integer")) ||
        !(word_ieq(t, tb[4], te[4], "symmetric") || word_ieq(t, tb[4], te[4], "skew-symmetric") ||
          word_ieq(t, tb[4], te[4], "general")))
        return HB_MTX_UNSUPPORTED;
    h->symmetric = !word_ieq(t, tb[4], te[4], "general");
    uint64_t b = e0 < len ? e0 + 1 : len, line = 1;
    for (;;) {
        if (b >= len) return HB_MTX_NO_SIZE_LINE;
        ++line;
        const char *q = (const char *)memchr(t + b, '\n', len - b);
        const uint64_t le = q ? (uint64_t)(q - t) : len;
        uint64_t lb = b, lt = le;
        b = le < len ? le + 1 : len;
        while (lb < lt && host_ws(t[lb])) ++lb;
        while (lt > lb && host_ws(t[lt - 1])) --lt;
        if (lb == lt || t[lb] == '%') continue;
        uint64_t v[3];
        const int k = tokens(t, lb, lt, tb, te, 3);
        for (int i = 0; i < 3; ++i)
            if (i >= k || !host_usize(t, tb[i], te[i], &v[i])) {
                h->err_line = line;
                return HB_MTX_BAD_SIZE_LINE;
            }
        if (v[0] != v[1]) {
            h->err_line = line;
            return HB_MTX_NOT_SQUARE;
        }
        h->n = v[0];
        h->nnz = v[2];
        h->body = b;
        h->body_line = line + 1;
        return 0;
    }
}

hipError_t mtx_body_to_csr(const uint8_t *text, uint64_t len, uint32_t n, bool symmetric, CsrBuild *out,
                           hipStream_t s)
{
    out->off = out->nbr = nullptr;
    out->arcs = out->emitted = out->entries = 0;
    out->err_index = ~0ull;
    out->err_code = 0;
    // line l = [starts[l], starts[l + 1]); starts[0] = 0, starts[lines] = len
    Tmp tmp;
    uint64_t nl = 0;
    Tmp starts;
    if (len > 1) {
        const uint64_t chunks = (len + kSplitChunk - 1) / kSplitChunk;
        Tmp per, base;
        HBI_TRY(per.alloc(chunks * 8));
        HBI_TRY(base.alloc(chunks * 8));
        k_count_nl<<<(unsigned)chunks, kBlock, 0, s>>>(text, len, (unsigned long long *)per.p);
        HBI_TRY(hipGetLastError());
        size_t bytes = 0;
        HBI_TRY(rocprim::exclusive_scan(nullptr, bytes, (const unsigned long long *)per.p,
                                        (unsigned long long *)base.p, 0ull, (size_t)chunks,
                                        rocprim::plus<unsigned long long>(), s));
        HBI_TRY(tmp.alloc(bytes));
        HBI_TRY(rocprim::exclusive_scan(tmp.p, bytes, (const unsigned long long *)per.p,
                                        (unsigned long long *)base.p, 0ull, (size_t)chunks,
                                        rocprim::plus<unsigned long long>(), s));
        unsigned long long last[2] = {0, 0};
        HBI_TRY(hipMemcpyAsync(&last[0], (unsigned long long *)base.p + chunks - 1, 8, hipMemcpyDeviceToHost, s));
        HBI_TRY(hipMemcpyAsync(&last[1], (unsigned long long *)per.p + chunks - 1, 8, hipMemcpyDeviceToHost, s));
        uint8_t tail = 0;
        HBI_TRY(hipMemcpyAsync(&tail, text + len - 1, 1, hipMemcpyDeviceToHost, s));
        HBI_TRY(hipStreamSynchronize(s));
        nl = last[0] + last[1] - (tail == '\n' ? 1u : 0u);  // a final '\n' starts no line
        HBI_TRY(starts.alloc((nl + 2) * 8));
        k_line_starts<<<(unsigned)chunks, kBlock, 0, s>>>(text, len, (const unsigned long long *)base.p,
                                                           (uint64_t *)starts.p);
        HBI_TRY(hipGetLastError());
    } else {
        HBI_TRY(starts.alloc(16));
    }
    const uint64_t lines = nl + 1;
    const uint64_t zero = 0;
    HBI_TRY(hipMemcpyAsync(starts.p, &zero, 8, hipMemcpyHostToDevice, s));
    HBI_TRY(hipMemcpyAsync((uint64_t *)starts.p + lines, &len, 8, hipMemcpyHostToDevice, s));
    Tmp emits, ei, ej, err;
    HBI_TRY(emits.alloc(lines * 4));
    HBI_TRY(ei.alloc(lines * 4));
    HBI_TRY(ej.alloc(lines * 4));
    HBI_TRY(err.alloc(8));
    HBI_TRY(hipMemsetAsync(err.p, 0xff, 8, s));
    k_parse_lines<<<grid_for(lines), kBlock, 0, s>>>(text, (const uint64_t *)starts.p, lines, n, symmetric ? 1 : 0,
                                                     (uint32_t *)emits.p, (uint32_t *)ei.p, (uint32_t *)ej.p,
                                                     (unsigned long long *)err.p);
    HBI_TRY(hipGetLastError());
    unsigned long long h_err = 0;
    HBI_TRY(hipMemcpyAsync(&h_err, err.p, 8, hipMemcpyDeviceToHost, s));
    HBI_TRY(hipStreamSynchronize(s));
    if (h_err != ~0ull) {
        out->err_index = h_err >> 4;
        out->err_code = (int)(h_err & 15u);
        return hipSuccess;
    }
    const hipError_t e = build_csr((const uint32_t *)ei.p, (const uint32_t *)ej.p, (const uint32_t *)emits.p, lines,
                                   n, false, out, s);
    if (e == hipSuccess) {
        // entries = lines that emitted
        Tmp c1, t1;
        HBI_TRY(c1.alloc(8));
        size_t rb = 0;
        auto is_entry = rocprim::make_transform_iterator((const uint32_t *)emits.p,
                                                         [] __device__(uint32_t x) { return (unsigned long long)(x != 0); });
        HBI_TRY(rocprim::reduce(nullptr, rb, is_entry, (unsigned long long *)c1.p, 0ull, (size_t)lines,
                                rocprim::plus<unsigned long long>(), s));
        HBI_TRY(t1.alloc(rb));
        HBI_TRY(rocprim::reduce(t1.p, rb, is_entry, (unsigned long long *)c1.p, 0ull, (size_t)lines,
                                rocprim::plus<unsigned long long>(), s));
        unsigned long long entries = 0;
        HBI_TRY(hipMemcpyAsync(&entries, c1.p, 8, hipMemcpyDeviceToHost, s));
        HBI_TRY(hipStreamSynchronize(s));
        out->entries = entries;
    }
    return e;
}

hipError_t coo_to_csr(const uint32_t *src, const uint32_t *dst, uint64_t m, uint32_t n, uint32_t flags,
                      CsrBuild *out, hipStream_t s)
{
    out->off = out->nbr = nullptr;
    out->arcs = out->emitted = out->entries = 0;
    out->err_index = ~0ull;
    out->err_code = 0;
    Tmp emits, err;
    HBI_TRY(emits.alloc(m * 4));
    HBI_TRY(err.alloc(8));
    HBI_TRY(hipMemsetAsync(err.p, 0xff, 8, s));
    if (m) {
        k_coo_emits<<<grid_for(m), kBlock, 0, s>>>(src, dst, m, n, flags, (uint32_t *)emits.p,
                                                   (unsigned long long *)err.p);
        HBI_TRY(hipGetLastError());
    }
    unsigned long long h_err = 0;
    HBI_TRY(hipMemcpyAsync(&h_err, err.p, 8, hipMemcpyDeviceToHost, s));
    HBI_TRY(hipStreamSynchronize(s));
    if (h_err != ~0ull) {
        out->err_index = h_err >> 4;
        out->err_code = (int)(h_err & 15u);
        return hipSuccess;
    }
    out->entries = m;
    return build_csr(src, dst, (const uint32_t *)emits.p, m, n, (flags & HB_COO_DEDUP) != 0, out, s);
}

}  // namespace ingest
}  // namespace hb
